"""CPU oracle for the qcmine GPU path (test infrastructure only).

ctypes wrapper of liboracle.so, a C++ restatement of the reference's host
expand/prune path (main.cu:846-2198), its RemoveNonMax (main.cu:4569-4961),
the N2 rule of adjToSer.cpp:115-156 and a brute-force enumerator.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .build import OUT, build

_lib = None

STAT_FIELDS = ["tasks", "branches", "children", "raw_cliques", "sum_t", "sum_t_written", "sum_clique_len", "w_adj",
               "lookahead_hits", "stale_oob", "stale_hits", "max_t", "max_depth", "root_size", "init_w_adj", "beyond",
               "root_branches",
               # how often each pruning rule fired
               "remove_breaks", "lu_low", "lu_lb_invalid", "lu_sum_invalid", "lu_ub_tightened", "lu_invalid",
               "member_fail", "dp_recompute", "dp_decrement", "crit_none", "crit_fail", "crit_ok"]


class _Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in STAT_FIELDS]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = build()
        L = C.CDLL(str(path))
        P = C.c_void_p
        L.orc_mine.restype = P
        L.orc_mine.argtypes = [C.c_int, P, P, P, P, C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_mine_window.restype = P
        L.orc_mine_window.argtypes = [C.c_int, P, P, P, P, C.c_double, C.c_int, C.c_int, C.c_int, C.c_long,
                                      C.c_long, C.c_int]
        L.orc_mine_window_mt.restype = P
        L.orc_mine_window_mt.argtypes = [C.c_int, P, P, P, P, C.c_double, C.c_int, C.c_int, C.c_int, C.c_long,
                                         C.c_long, C.c_int, C.c_int]
        L.orc_mine_window2.restype = P
        L.orc_mine_window2.argtypes = [C.c_int, P, P, P, P, C.c_double, C.c_int, C.c_int, C.c_int, C.c_long,
                                       C.c_long, C.c_long, C.c_long, C.c_int, C.c_int]
        L.orc_count.restype = C.c_uint64
        L.orc_count.argtypes = [P]
        L.orc_off.restype = C.POINTER(C.c_uint64)
        L.orc_off.argtypes = [P]
        L.orc_verts.restype = C.POINTER(C.c_int)
        L.orc_verts.argtypes = [P]
        L.orc_get_stats.argtypes = [P, C.POINTER(_Stats)]
        L.orc_trace_rows.restype = C.c_uint64
        L.orc_trace_rows.argtypes = [P]
        L.orc_trace_feats.restype = C.POINTER(C.c_float)
        L.orc_trace_feats.argtypes = [P]
        L.orc_trace_parent.restype = C.POINTER(C.c_int64)
        L.orc_trace_parent.argtypes = [P]
        L.orc_trace_beyond.restype = C.POINTER(C.c_int64)
        L.orc_trace_beyond.argtypes = [P]
        L.orc_free.argtypes = [P]
        L.orc_nonmax.restype = P
        L.orc_nonmax.argtypes = [C.c_uint64, P, P]
        L.orc_build_n2.restype = C.c_uint64
        L.orc_build_n2.argtypes = [C.c_int, P, P, P, P]
        L.orc_lu_bounds.restype = None
        L.orc_lu_bounds.argtypes = [C.c_double, C.c_int, C.c_int, P, P, C.c_int, P, P]
        L.orc_root.restype = C.c_int
        L.orc_root.argtypes = [C.c_int, P, P, C.c_double, C.c_int, P, P, P]
        L.orc_brute.restype = P
        L.orc_brute.argtypes = [C.c_int, P, P, C.c_double, C.c_int]
        _lib = L
    return _lib


def _take(h) -> list[np.ndarray]:
    L = lib()
    k = L.orc_count(h)
    if k == 0:
        return []
    off = np.ctypeslib.as_array(L.orc_off(h), shape=(k + 1,)).copy()
    v = np.ctypeslib.as_array(L.orc_verts(h), shape=(int(off[-1]),)).copy() if off[-1] else np.zeros(0, np.int32)
    return [v[off[i]:off[i + 1]].astype(np.int64) for i in range(k)]


def _csr(off, adj):
    off = np.ascontiguousarray(off, dtype=np.uint64)
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    return off, adj


def mine(off, adj, gamma: float, min_size: int, *, faithful: bool = False, trace: bool = False, off2=None,
         adj2=None, root_begin: int = 0, root_end: int = -1, max_depth: int = -1, threads: int = 1,
         l1_begin: int = 0, l1_end: int = -1) -> dict:
    """Run the restated host path to completion (or over a search window:
    root branches [root_begin, root_end), branches [l1_begin, l1_end) of every
    depth-1 task and depth <= max_depth); returns raw sets, stats and
    (optionally) the per-task trace.  threads > 1 runs the
    same search on that many host threads (clean map mode, no trace): equal
    counters and raw multiset, raw sets in another order."""
    if threads > 1 and (faithful or trace):
        raise ValueError("the multi-threaded search runs in clean map mode without a trace")
    L = lib()
    off, adj = _csr(off, adj)
    o2 = a2 = None
    if off2 is not None:
        off2, adj2 = _csr(off2, adj2)
        o2, a2 = off2.ctypes.data, adj2.ctypes.data
    h = L.orc_mine_window2(len(off) - 1, off.ctypes.data, adj.ctypes.data if adj.size else None, o2, a2,
                           float(gamma), int(min_size), int(faithful), int(trace), int(root_begin), int(root_end),
                           int(l1_begin), int(l1_end), int(max_depth), int(threads))
    try:
        st = _Stats()
        L.orc_get_stats(h, C.byref(st))
        out = {"raw": _take(h), "stats": {n: int(getattr(st, n)) for n in STAT_FIELDS}}
        if trace:
            n = L.orc_trace_rows(h)
            out["trace_feats"] = np.ctypeslib.as_array(L.orc_trace_feats(h), shape=(n, 12)).copy() if n else \
                np.zeros((0, 12), np.float32)
            out["trace_parent"] = np.ctypeslib.as_array(L.orc_trace_parent(h), shape=(n,)).copy() if n else \
                np.zeros(0, np.int64)
            out["trace_beyond"] = np.ctypeslib.as_array(L.orc_trace_beyond(h), shape=(n,)).copy() if n else \
                np.zeros(0, np.int64)
        return out
    finally:
        L.orc_free(h)


def nonmax(sets) -> list[tuple[int, ...]]:
    """RemoveNonMax semantics: distinct sets not strictly contained in another,
    each ascending, lexicographic order."""
    L = lib()
    sets = [np.asarray(s, np.int32) for s in sets]
    off = np.zeros(len(sets) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in sets]) if sets else []
    v = np.concatenate(sets).astype(np.int32) if sets else np.zeros(0, np.int32)
    h = L.orc_nonmax(len(sets), off.ctypes.data, v.ctypes.data if v.size else None)
    try:
        return [tuple(int(x) for x in s) for s in _take(h)]
    finally:
        L.orc_free(h)


def brute(off, adj, gamma: float, min_size: int) -> list[tuple[int, ...]]:
    """All maximal gamma-quasi-cliques of size >= min_size by exhaustive search (n <= 24)."""
    L = lib()
    off, adj = _csr(off, adj)
    assert len(off) - 1 <= 24
    h = L.orc_brute(len(off) - 1, off.ctypes.data, adj.ctypes.data if adj.size else None, float(gamma), int(min_size))
    try:
        return [tuple(int(x) for x in s) for s in _take(h)]
    finally:
        L.orc_free(h)


def build_n2(off, adj) -> tuple[np.ndarray, np.ndarray]:
    """Full-graph N2 lists (adjToSer.cpp:115-156)."""
    L = lib()
    off, adj = _csr(off, adj)
    n = len(off) - 1
    off2 = np.zeros(n + 1, np.uint64)
    tot = L.orc_build_n2(n, off.ctypes.data, adj.ctypes.data if adj.size else None, off2.ctypes.data, None)
    adj2 = np.zeros(max(int(tot), 1), np.int32)
    L.orc_build_n2(n, off.ctypes.data, adj.ctypes.data if adj.size else None, off2.ctypes.data, adj2.ctypes.data)
    return off2, adj2[:tot]


def root(off, adj, gamma: float, min_size: int) -> dict:
    """The root task record (initialize_tasks, main.cu:846-998) in its Q
    order: vertex ids, labels (0 candidate, 2 covered, 3 cover), exdegs."""
    L = lib()
    off, adj = _csr(off, adj)
    n = len(off) - 1
    ids, labels, ex = (np.zeros(max(n, 1), np.int32) for _ in range(3))
    k = L.orc_root(n, off.ctypes.data, adj.ctypes.data if adj.size else None, float(gamma), int(min_size),
                   ids.ctypes.data, labels.ctypes.data, ex.ctypes.data)
    return {"ids": ids[:k].copy(), "labels": labels[:k].copy(), "exdeg": ex[:k].copy()}


def lu_bounds(gamma: float, min_size: int, members, cand_indeg) -> tuple[bool, int, int, int]:
    """The oracle's LU bounds (main.cu:1857-1972) for members given as
    (indeg, exdeg) pairs and candidate indegrees (sorted descending here):
    (invalid, lb, ub, min_ext_deg)."""
    L = lib()
    ind = np.array([x[0] for x in members], np.int32)
    exd = np.array([x[1] for x in members], np.int32)
    ci = np.array(sorted(cand_indeg, reverse=True) or [0], np.int32)
    out = np.zeros(4, np.int32)
    L.orc_lu_bounds(float(gamma), int(min_size), len(members), ind.ctypes.data, exd.ctypes.data, len(cand_indeg),
                    ci.ctypes.data, out.ctypes.data)
    return bool(out[0]), int(out[1]), int(out[2]), int(out[3])


def canon(sets) -> list[tuple[int, ...]]:
    """Multiset of sets as a sorted list of ascending tuples."""
    return sorted(tuple(sorted(int(x) for x in s)) for s in sets)
