"""The serialized writer with N2 built on the GPU (qc_graph_write_serialized_device,
adjToSer.cpp:115-200): every file must equal, byte for byte, the host
writer's file for the same graph (whose N2 rule is pinned on the CPU against
the oracle's N2 builder).  Cases: generated graphs, a hub whose neighbours'
lists overlap heavily, isolated vertices, n not a multiple of 64, a million
sparse vertices (a 125 KB LDS map), many small output chunks, and a graph too
large for the LDS map (an error before any launch)."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

import cuqc_amd as qc
from cuqc_amd import workloads

pytestmark = pytest.mark.gpu

CONV = str(qc.library_path().parent / "qcconv")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _csr(n, edges):
    """Symmetric CSR from undirected pairs (no self loops, no duplicates)."""
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    e = e[e[:, 0] != e[:, 1]]
    both = np.concatenate([e, e[:, ::-1]])
    both = np.unique(both, axis=0)
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, both[:, 0] + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    return qc.Graph.from_csr(off, both[:, 1].astype(np.uint32))


def _hub_graph():
    rng = np.random.default_rng(7)
    n = 5000
    edges = [(0, int(v)) for v in rng.choice(np.arange(1, n), 3000, replace=False)]
    edges += [tuple(map(int, rng.integers(0, n, 2))) for _ in range(20000)]
    # vertices 4990..4999 left isolated
    return _csr(n, [(u, v) for u, v in edges if u < 4990 and v < 4990])


def _million():
    rng = np.random.default_rng(3)
    n = 1_000_000
    perm = rng.permutation(n)
    edges = np.stack([perm[0::2], perm[1::2]], 1)          # a perfect matching
    path = np.arange(0, 200_000, dtype=np.int64)
    edges = np.concatenate([edges, np.stack([path[:-1], path[1:]], 1)])
    return _csr(n, edges)


CASES = {
    "config0": lambda: qc.Graph.generate(0, 1),
    "batch_instance12": lambda: workloads.graph(workloads.WORKLOADS[3], 12),
    "hub_isolated": _hub_graph,
    "n70": lambda: _csr(70, [(i, (i * 7 + 3) % 70) for i in range(70)]),
    "single_vertex": lambda: _csr(1, []),
    "empty": lambda: _csr(0, []),
    "million_sparse": _million,
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_file_equals_host_file(tmp_path, name):
    g = CASES[name]()
    _say(f"[gpu] n2 writer {name}: n {g.n}")
    host, dev = tmp_path / "host.txt", tmp_path / "dev.txt"
    g.write_serialized(host)
    g.write_serialized(dev, device=True)
    assert dev.read_bytes() == host.read_bytes(), name
    if g.n:
        back = qc.Graph.load_serialized(dev)
        assert back.n == g.n


def test_many_chunks(tmp_path, monkeypatch):
    g = _hub_graph()
    host, dev = tmp_path / "host.txt", tmp_path / "dev.txt"
    g.write_serialized(host)
    # chunks of at most n entries: the hub's neighbours' lists span many
    monkeypatch.setenv("QCMINE_N2_CHUNK", "1")
    _say("[gpu] n2 writer: chunks of n entries")
    g.write_serialized(dev, device=True)
    assert dev.read_bytes() == host.read_bytes()


def test_too_many_vertices_is_an_error(tmp_path):
    g = _csr(2_000_000, [(0, 1)])
    _say("[gpu] n2 writer: 2M vertices rejected")
    with pytest.raises(qc.QCError) as e:
        g.write_serialized(tmp_path / "x.txt", device=True)
    assert e.value.code == qc.QC_EINVAL


@pytest.mark.skip(reason="qcconv commands that called the device writer from a subprocess (adj-to-ser-gpu, "
                         "bin-to-ser) stalled once, with no cause found in the code; the commands were taken "
                         "out and the writer is reached through the library call only")
def test_qcconv_gpu_commands(tmp_path):
    g = workloads.graph(workloads.WORKLOADS[3], 12)
    off, adj = g.csr()
    adjf = tmp_path / "g.adj"
    adjf.write_text("".join(" ".join(str(int(x)) for x in adj[off[v]:off[v + 1]]) + "\n" for v in range(g.n)))
    binf = tmp_path / "g.bin"
    g.write_binary(binf)
    ref, a, b = tmp_path / "ref.txt", tmp_path / "a.txt", tmp_path / "b.txt"
    _say("[gpu] qcconv adj-to-ser / adj-to-ser-gpu / bin-to-ser")
    for cmd, src, dst in (("adj-to-ser", adjf, ref), ("adj-to-ser-gpu", adjf, a), ("bin-to-ser", binf, b)):
        r = subprocess.run([CONV, cmd, str(src), str(dst)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (cmd, r.stdout, r.stderr)
    assert a.read_bytes() == ref.read_bytes()
    assert b.read_bytes() == ref.read_bytes()
