"""cuqc_amd -- maximal gamma-quasi-clique mining on AMD MI355X (gfx950).

Python bindings (ctypes) of ``libqcmine.so``, whose C ABI is declared in
``include/qcmine.h``.  The reference (Mike12041204/cuQC) is driven by its
command line ``cuQC graph_file gamma min_size output_file scheduling_toggle``
(main.cu:416-519); :func:`mine` takes the same arguments, and the ``qcmine``
binary next to this file keeps that command line.

The library needs a gfx950 device: every mining call raises :class:`QCError`
when none is present.  There is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
from collections.abc import Sequence
from dataclasses import dataclass
from pathlib import Path

import numpy as np

__all__ = [
    "QCError", "Graph", "Result", "CostModel", "mine", "mine_batch", "remove_nonmax", "device_available",
    "library_path", "load_library",
]

_PKG = Path(__file__).resolve().parent
_LIBPATH = _PKG / "libqcmine.so"

QC_OK, QC_EINVAL, QC_EIO, QC_ENOMEM, QC_EOVERFLOW, QC_EDEVICE, QC_EEMPTY = range(7)
_MODES = {"plain": 0, "persistent": 1}
_ORDERS = {"fifo": 0, "cost": 1}


class QCError(RuntimeError):
    """A library call failed; ``code`` is the qc_status value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[qc_status {code}] {msg}")
        self.code = code


class _Params(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("min_size", C.c_int32), ("scheduling", C.c_int32), ("mode", C.c_int32),
        ("order", C.c_int32), ("workgroups", C.c_int32), ("lds_tile", C.c_int32), ("split_chunk", C.c_int32),
        ("max_depth", C.c_int32), ("task_capacity", C.c_uint64), ("clique_capacity", C.c_uint64),
        ("cost_model", C.c_void_p), ("root_begin", C.c_int64), ("root_end", C.c_int64),
        ("replays", C.c_int32), ("trace", C.c_int32), ("split_min", C.c_int32), ("mid_tile", C.c_int32),
        ("lds_wide", C.c_int32),
    ]


_STAT_U64 = ["tasks", "branches", "children", "raw_cliques", "max_cliques", "sum_t", "sum_t_written",
             "sum_clique_len", "root_size", "levels", "large_tasks", "beyond", "root_branches", "w_adj"]


class _Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in _STAT_U64] + [("pool_peak", C.c_uint64), ("reserved", C.c_uint64)] + [
        ("t_init_ms", C.c_double), ("t_enum_ms", C.c_double), ("t_nonmax_ms", C.c_double), ("t_kernel_ms", C.c_double)]


_lib = None


def library_path() -> Path:
    return _LIBPATH


def load_library() -> C.CDLL:
    """Load libqcmine.so (built in-tree by ``python -m cuqc_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIBPATH.exists():
        raise QCError(QC_EDEVICE, f"{_LIBPATH} is not built; run `python -m cuqc_amd.build`")
    L = C.CDLL(str(_LIBPATH))
    P, PP = C.c_void_p, C.POINTER(C.c_void_p)
    u64p = C.POINTER(C.c_uint64)
    sig = {
        "qc_graph_load_serialized": (C.c_int, [C.c_char_p, PP]),
        "qc_graph_load_adjlist": (C.c_int, [C.c_char_p, PP]),
        "qc_graph_write_serialized": (C.c_int, [P, C.c_char_p]),
        "qc_graph_write_serialized_device": (C.c_int, [P, C.c_char_p]),
        "qc_graph_load_binary": (C.c_int, [C.c_char_p, PP]),
        "qc_graph_write_binary": (C.c_int, [P, C.c_char_p]),
        "qc_graph_from_csr": (C.c_int, [C.c_int32, P, P, PP]),
        "qc_graph_generate": (C.c_int, [C.c_int32, C.c_uint64, C.c_int32, PP, C.POINTER(C.c_double),
                                        C.POINTER(C.c_int32)]),
        "qc_graph_csr": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(u64p), C.POINTER(C.POINTER(C.c_uint32))]),
        "qc_graph_checksum": (C.c_uint64, [P]),
        "qc_graph_planted": (C.c_int, [P, u64p, C.POINTER(u64p), C.POINTER(C.POINTER(C.c_uint32))]),
        "qc_graph_free": (None, [P]),
        "qc_mine": (C.c_int, [P, C.POINTER(_Params), PP]),
        "qc_mine_batch": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(_Params), C.c_int32, PP]),
        "qc_result_view": (C.c_int, [P, u64p, C.POINTER(u64p), C.POINTER(C.POINTER(C.c_uint32))]),
        "qc_result_raw": (C.c_int, [P, u64p, C.POINTER(u64p), C.POINTER(C.POINTER(C.c_uint32))]),
        "qc_result_stats": (C.c_int, [P, C.POINTER(_Stats)]),
        "qc_result_trace": (C.c_int, [P, u64p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.POINTER(C.c_int64)),
                                      C.POINTER(C.POINTER(C.c_int64))]),
        "qc_convert_edge_to_adj": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int32]),
        "qc_convert_rm_empty": (C.c_int, [C.c_char_p, C.c_char_p]),
        "qc_convert_adj_to_ser": (C.c_int, [C.c_char_p, C.c_char_p]),
        "qc_result_write": (C.c_int, [P, C.c_char_p]),
        "qc_result_write_raw": (C.c_int, [P, C.c_char_p]),
        "qc_result_free": (None, [P]),
        "qc_remove_nonmax": (C.c_int, [C.c_uint64, P, P, PP]),
        "qc_degree_tables": (C.c_int, [C.c_double, C.c_int32, P, C.c_int32, P]),
        "qc_model_train": (C.c_int, [P, P, C.c_uint64, C.c_int32, C.c_float, C.c_uint64, PP]),
        "qc_model_predict": (C.c_int, [P, P, C.c_uint64, P]),
        "qc_model_grad": (C.c_int, [P, P, P, C.c_uint64, P, P]),
        "qc_model_create": (C.c_int, [C.c_uint64, PP]),
        "qc_model_params": (C.c_int, [P, P, P, P, P]),
        "qc_model_save": (C.c_int, [P, C.c_char_p]),
        "qc_model_load": (C.c_int, [C.c_char_p, PP]),
        "qc_model_free": (None, [P]),
        "qc_last_error": (C.c_char_p, []),
        "qc_version": (C.c_char_p, []),
        "qc_device_available": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int, ok=(QC_OK,)) -> int:
    if rc not in ok:
        raise QCError(rc, load_library().qc_last_error().decode(errors="replace"))
    return rc


def device_available() -> bool:
    return bool(load_library().qc_device_available())


def _sets_from(count: int, off, verts) -> list[np.ndarray]:
    if count == 0:
        return []
    o = np.ctypeslib.as_array(off, shape=(count + 1,)).copy()
    v = np.ctypeslib.as_array(verts, shape=(int(o[-1]),)).copy() if o[-1] else np.zeros(0, np.uint32)
    return [v[o[i]:o[i + 1]] for i in range(count)]


class SetList(Sequence):
    """Vertex sets held as one offsets array (count + 1 entries) and one
    vertex array; indexing and iteration give each set as an array view, so
    millions of raw sets cost two arrays, not a list of millions."""

    __slots__ = ("offsets", "verts")

    def __init__(self, offsets: np.ndarray, verts: np.ndarray):
        self.offsets = offsets
        self.verts = verts

    @classmethod
    def from_view(cls, count: int, off, verts) -> "SetList":
        if count == 0:
            return cls(np.zeros(1, np.uint64), np.zeros(0, np.uint32))
        o = np.ctypeslib.as_array(off, shape=(count + 1,)).copy()
        v = np.ctypeslib.as_array(verts, shape=(int(o[-1]),)).copy() if o[-1] else np.zeros(0, np.uint32)
        return cls(o, v)

    def __len__(self) -> int:
        return len(self.offsets) - 1

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(len(self)))]
        n = len(self)
        if i < 0:
            i += n
        if not 0 <= i < n:
            raise IndexError(i)
        return self.verts[self.offsets[i]:self.offsets[i + 1]]

    def __iter__(self):
        o, v = self.offsets, self.verts
        for i in range(len(self)):
            yield v[o[i]:o[i + 1]]


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def multiset_hash(sets: "SetList") -> str:
    """Hash of a multiset of vertex sets that ignores the order of the sets
    and of the vertices in each (a fast equality check between runs, not a
    cryptographic digest): per set, the sum of a 64-bit mix of every vertex,
    mixed with the set's size; the sorted per-set values are hashed."""
    o = np.asarray(sets.offsets, np.int64)
    n = len(o) - 1
    if n <= 0:
        return hashlib.sha256(b"").hexdigest()
    lens = np.diff(o).astype(np.uint64)
    with np.errstate(over="ignore"):
        x = _mix64(np.asarray(sets.verts, np.uint64))
        cs = np.concatenate(([np.uint64(0)], np.cumsum(x, dtype=np.uint64)))
        per = _mix64((cs[o[1:]] - cs[o[:-1]]) ^ _mix64(lens))
    per.sort()
    return hashlib.sha256(per.tobytes()).hexdigest()


class Graph:
    """An undirected graph held by the library (symmetric CSR, sorted lists)."""

    def __init__(self, handle: int, gamma: float | None = None, min_size: int | None = None):
        self._h = C.c_void_p(handle)
        self.gamma = gamma
        self.min_size = min_size

    # -- constructors ---------------------------------------------------
    @classmethod
    def from_csr(cls, offsets, adj) -> "Graph":
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        a = np.ascontiguousarray(adj, dtype=np.uint32)
        h = C.c_void_p()
        _check(load_library().qc_graph_from_csr(len(off) - 1, off.ctypes.data, a.ctypes.data if a.size else None,
                                                C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges(cls, n: int, edges) -> "Graph":
        e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        e = e[e[:, 0] != e[:, 1]]
        both = np.concatenate([e, e[:, ::-1]])
        both = np.unique(both, axis=0) if len(both) else both.reshape(0, 2)
        off = np.zeros(n + 1, np.uint64)
        if len(both):
            np.add.at(off, both[:, 0] + 1, 1)
        off = np.cumsum(off).astype(np.uint64)
        return cls.from_csr(off, both[:, 1].astype(np.uint32) if len(both) else np.zeros(0, np.uint32))

    @classmethod
    def generate(cls, config: int, seed: int, instance: int = 0) -> "Graph":
        h, g, m = C.c_void_p(), C.c_double(), C.c_int32()
        _check(load_library().qc_graph_generate(config, seed, instance, C.byref(h), C.byref(g), C.byref(m)))
        return cls(h.value, g.value, m.value)

    @classmethod
    def load_serialized(cls, path: str | os.PathLike) -> "Graph":
        h = C.c_void_p()
        _check(load_library().qc_graph_load_serialized(os.fsencode(path), C.byref(h)))
        return cls(h.value)

    @classmethod
    def load_adjlist(cls, path: str | os.PathLike) -> "Graph":
        h = C.c_void_p()
        _check(load_library().qc_graph_load_adjlist(os.fsencode(path), C.byref(h)))
        return cls(h.value)

    @classmethod
    def load_binary(cls, path: str | os.PathLike) -> "Graph":
        h = C.c_void_p()
        _check(load_library().qc_graph_load_binary(os.fsencode(path), C.byref(h)))
        return cls(h.value)

    # -- views ----------------------------------------------------------
    def csr(self) -> tuple[np.ndarray, np.ndarray]:
        n = C.c_int32()
        o = C.POINTER(C.c_uint64)()
        a = C.POINTER(C.c_uint32)()
        _check(load_library().qc_graph_csr(self._h, C.byref(n), C.byref(o), C.byref(a)))
        off = np.ctypeslib.as_array(o, shape=(n.value + 1,)).copy()
        adj = np.ctypeslib.as_array(a, shape=(int(off[-1]),)).copy() if off[-1] else np.zeros(0, np.uint32)
        return off, adj

    @property
    def n(self) -> int:
        n = C.c_int32()
        _check(load_library().qc_graph_csr(self._h, C.byref(n), None, None))
        return n.value

    def checksum(self) -> int:
        return int(load_library().qc_graph_checksum(self._h))

    def planted(self) -> list[np.ndarray]:
        """Vertex sets of the quasi-cliques a generated graph was planted with."""
        cnt, off, v = C.c_uint64(), C.POINTER(C.c_uint64)(), C.POINTER(C.c_uint32)()
        _check(load_library().qc_graph_planted(self._h, C.byref(cnt), C.byref(off), C.byref(v)))
        return _sets_from(cnt.value, off, v)

    def write_serialized(self, path, device: bool = False) -> None:
        """The reference's serialized text format (adjToSer.cpp).  device:
        N2 counted and built on the GPU, written in chunks (large graphs)."""
        L = load_library()
        fn = L.qc_graph_write_serialized_device if device else L.qc_graph_write_serialized
        _check(fn(self._h, os.fsencode(path)))

    def write_binary(self, path) -> None:
        _check(load_library().qc_graph_write_binary(self._h, os.fsencode(path)))

    def close(self) -> None:
        if self._h:
            load_library().qc_graph_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Result:
    """Maximal sets (ascending, lexicographic order), raw sets and statistics."""

    maximal: list[np.ndarray]
    raw: SetList
    stats: dict
    status: int
    trace_feats: np.ndarray | None = None     # (rows, 12) float32
    trace_parent: np.ndarray | None = None    # row of the parent task, -1 root
    trace_subtree: np.ndarray | None = None   # tasks in the subtree

    def write(self, path) -> None:
        write_sets(self.maximal, path)

    def maximal_sorted_tuples(self) -> list[tuple[int, ...]]:
        return [tuple(int(x) for x in s) for s in self.maximal]


def write_sets(sets: Sequence[Sequence[int]], path) -> None:
    """OutputOneSet format (main.cu:4641-4651)."""
    with open(path, "w") as f:
        for s in sets:
            f.write(f"{len(s)} " + "".join(f"{int(x)} " for x in s) + "\n")


def _result_from(h: C.c_void_p, rc: int) -> Result:
    L = load_library()
    try:
        cnt, off, v = C.c_uint64(), C.POINTER(C.c_uint64)(), C.POINTER(C.c_uint32)()
        _check(L.qc_result_view(h, C.byref(cnt), C.byref(off), C.byref(v)))
        maximal = _sets_from(cnt.value, off, v)
        _check(L.qc_result_raw(h, C.byref(cnt), C.byref(off), C.byref(v)))
        raw = SetList.from_view(cnt.value, off, v)
        st = _Stats()
        _check(L.qc_result_stats(h, C.byref(st)))
        stats = {n: int(getattr(st, n)) for n in _STAT_U64}
        stats.update(pool_peak=int(st.pool_peak), t_init_ms=st.t_init_ms, t_enum_ms=st.t_enum_ms,
                     t_nonmax_ms=st.t_nonmax_ms, t_kernel_ms=st.t_kernel_ms)
        rows = C.c_uint64()
        fp, pp, sp = C.POINTER(C.c_float)(), C.POINTER(C.c_int64)(), C.POINTER(C.c_int64)()
        _check(L.qc_result_trace(h, C.byref(rows), C.byref(fp), C.byref(pp), C.byref(sp)))
        tf = tpar = tsub = None
        if rows.value:
            n = rows.value
            tf = np.ctypeslib.as_array(fp, shape=(n, 12)).copy()
            tpar = np.ctypeslib.as_array(pp, shape=(n,)).copy()
            tsub = np.ctypeslib.as_array(sp, shape=(n,)).copy()
        return Result(maximal, raw, stats, rc, tf, tpar, tsub)
    finally:
        L.qc_result_free(h)


def _params(gamma, min_size, scheduling, mode, order, lds_tile, workgroups, task_capacity, clique_capacity,
            cost_model=None, split_chunk=0, window=None, replays=1, trace=False, split_min=0,
            mid_tile=0, lds_wide=0) -> _Params:
    p = _Params()
    p.gamma = float(gamma)
    p.min_size = int(min_size)
    p.scheduling = int(scheduling)
    p.mode = _MODES[mode] if isinstance(mode, str) else int(mode)
    p.order = _ORDERS[order] if isinstance(order, str) else int(order)
    p.lds_tile = int(lds_tile)
    p.workgroups = int(workgroups)
    p.split_chunk = int(split_chunk)
    p.task_capacity = int(task_capacity)
    p.clique_capacity = int(clique_capacity)
    p.cost_model = cost_model.handle if isinstance(cost_model, CostModel) else cost_model
    rb, re_, md = window if window is not None else (0, -1, -1)
    p.root_begin, p.root_end, p.max_depth = int(rb), max(0, int(re_)), max(0, int(md))
    p.replays = int(replays)
    p.trace = int(bool(trace))
    p.split_min = int(split_min)
    p.mid_tile = int(mid_tile)
    p.lds_wide = int(lds_wide)
    return p


def mine(graph: Graph, gamma: float, min_size: int, scheduling: int = 0, *, mode: str = "plain",
         order: str = "fifo", lds_tile: int = 0, workgroups: int = 0, task_capacity: int = 0,
         clique_capacity: int = 0, cost_model=None, split_chunk: int = 0, window=None, replays: int = 1,
         trace: bool = False, split_min: int = 0, mid_tile: int = 0, lds_wide: int = 0) -> Result:
    """Mine maximal gamma-quasi-cliques of size >= min_size (search() + RemoveNonMax()).

    ``window=(root_begin, root_end, max_depth)`` restricts the search to root
    branches [root_begin, root_end) and to tasks of depth <= max_depth
    (-1 = unbounded).  An empty result is returned (status QC_EEMPTY) rather
    than raised.
    """
    p = _params(gamma, min_size, scheduling, mode, order, lds_tile, workgroups, task_capacity, clique_capacity,
                cost_model, split_chunk, window, replays, trace, split_min, mid_tile, lds_wide)
    h = C.c_void_p()
    rc = _check(load_library().qc_mine(graph._h, C.byref(p), C.byref(h)), ok=(QC_OK, QC_EEMPTY))
    return _result_from(h, rc)


def mine_batch(graphs: Sequence[Graph], gammas: Sequence[float], min_sizes: Sequence[int], *, scheduling: int = 0,
               order: str = "fifo", lds_tile: int = 0, workgroups: int = 0, task_capacity: int = 0,
               clique_capacity: int = 0, cost_model=None, split_chunk: int = 0, windows=None,
               replays: int = 1, trace: bool = False, split_min: int = 0, mid_tile: int = 0,
               lds_wide: int = 0) -> list[Result]:
    """Many instances through one persistent launch."""
    n = len(graphs)
    ps = (_Params * n)()
    for i in range(n):
        ps[i] = _params(gammas[i], min_sizes[i], scheduling, "persistent", order, lds_tile, workgroups,
                        task_capacity, clique_capacity, cost_model, split_chunk,
                        windows[i] if windows is not None else None, replays, trace, split_min, mid_tile,
                        lds_wide)
    hs = (C.c_void_p * n)(*[g._h for g in graphs])
    out = (C.c_void_p * n)()
    _check(load_library().qc_mine_batch(hs, ps, n, out))
    return [_result_from(C.c_void_p(out[i]), QC_OK) for i in range(n)]


NPARAMS = 5057
NNORM = 29


class CostModel:
    """Task-cost MLP (12 -> 64 -> 64 -> 1, bf16 weights, fp32 accumulation)
    trained and evaluated by the library's GPU kernels."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @classmethod
    def create(cls, seed: int = 0) -> "CostModel":
        h = C.c_void_p()
        _check(load_library().qc_model_create(seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def train(cls, feats, log_target, steps: int = 300, lr: float = 3e-3, seed: int = 0) -> "CostModel":
        X = np.ascontiguousarray(feats, np.float32)
        y = np.ascontiguousarray(log_target, np.float32)
        h = C.c_void_p()
        _check(load_library().qc_model_train(X.ctypes.data, y.ctypes.data, len(y), steps, lr, seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, path) -> "CostModel":
        h = C.c_void_p()
        _check(load_library().qc_model_load(os.fsencode(path), C.byref(h)))
        return cls(h.value)

    def save(self, path) -> None:
        _check(load_library().qc_model_save(self._h, os.fsencode(path)))

    def predict(self, feats) -> np.ndarray:
        X = np.ascontiguousarray(feats, np.float32)
        out = np.zeros(len(X), np.float32)
        _check(load_library().qc_model_predict(self._h, X.ctypes.data, len(X), out.ctypes.data))
        return out

    def grad(self, feats, log_target) -> tuple[np.ndarray, float]:
        X = np.ascontiguousarray(feats, np.float32)
        y = np.ascontiguousarray(log_target, np.float32)
        g = np.zeros(NPARAMS, np.float32)
        loss = C.c_float()
        _check(load_library().qc_model_grad(self._h, X.ctypes.data, y.ctypes.data, len(y), g.ctypes.data,
                                            C.byref(loss)))
        return g, loss.value

    def params(self) -> tuple[np.ndarray, np.ndarray]:
        p = np.zeros(NPARAMS, np.float32)
        nrm = np.zeros(NNORM, np.float32)
        _check(load_library().qc_model_params(self._h, p.ctypes.data, None, nrm.ctypes.data, None))
        return p, nrm

    def set_params(self, p=None, norm=None) -> None:
        pp = np.ascontiguousarray(p, np.float32) if p is not None else None
        nn = np.ascontiguousarray(norm, np.float32) if norm is not None else None
        _check(load_library().qc_model_params(self._h, None, pp.ctypes.data if pp is not None else None, None,
                                              nn.ctypes.data if nn is not None else None))

    @property
    def handle(self):
        return self._h.value

    def __del__(self):
        try:
            if self._h:
                load_library().qc_model_free(self._h)
        except Exception:
            pass


def convert_edge_to_adj(in_path, out_path, duplicate: bool = False) -> None:
    """edgeToAdj.py: edge list -> adjacency list text."""
    _check(load_library().qc_convert_edge_to_adj(os.fsencode(in_path), os.fsencode(out_path), int(bool(duplicate))))


def convert_rm_empty(in_path, out_path) -> None:
    """rmEmptyVert.py: drop empty adjacency lines and renumber."""
    _check(load_library().qc_convert_rm_empty(os.fsencode(in_path), os.fsencode(out_path)))


def convert_adj_to_ser(in_path, out_path) -> None:
    """adjToSer.cpp: adjacency list text -> serialized graph with N2 lists."""
    _check(load_library().qc_convert_adj_to_ser(os.fsencode(in_path), os.fsencode(out_path)))


def degree_tables(gamma: float, n: int) -> tuple[np.ndarray, np.ndarray]:
    """mindeg[0..n) and ubq[0..n) as the library builds them."""
    md = np.zeros(n, np.int32)
    uq = np.zeros(n, np.int32)
    _check(load_library().qc_degree_tables(float(gamma), n, md.ctypes.data, n, uq.ctypes.data))
    return md, uq


def remove_nonmax(sets: Sequence[Sequence[int]]) -> list[np.ndarray]:
    """RemoveNonMax on the device for caller-provided raw sets."""
    lens = np.array([len(s) for s in sets], np.uint64)
    off = np.zeros(len(sets) + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    v = np.concatenate([np.asarray(s, np.uint32) for s in sets]) if len(sets) else np.zeros(0, np.uint32)
    h = C.c_void_p()
    rc = _check(load_library().qc_remove_nonmax(len(sets), off.ctypes.data, v.ctypes.data if v.size else None,
                                                C.byref(h)), ok=(QC_OK, QC_EEMPTY))
    return _result_from(h, rc).maximal
