"""C ABI, graph formats, generators and the CLI surface (CPU only: no call
here reaches the GPU)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import cuqc_amd as qc
import oracle
from tests.graphs import planted

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qcmine.h")
CLI = os.path.join(ROOT, "cuqc_amd", "qcmine")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = qc.load_library()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in qc.load_library().qc_version()


GEN_CHECKSUMS = {
    (0, 1, 0): 0xc1b55b3e7f27478c,
    (1, 2, 0): 0xc2237dec155d05f3,
}


@pytest.mark.parametrize("key", sorted(GEN_CHECKSUMS))
def test_generator_checksums(key):
    g = qc.Graph.generate(*key)
    assert g.checksum() == GEN_CHECKSUMS[key]


def test_generator_properties():
    g = qc.Graph.generate(0, 1)
    off, adj = g.csr()
    n = g.n
    assert n == 200 and g.gamma == 0.9 and g.min_size == 10
    # symmetric, sorted, no self loops
    for v in range(n):
        lst = adj[off[v]:off[v + 1]]
        assert np.all(np.diff(lst.astype(np.int64)) > 0)
        assert v not in lst
        for u in lst:
            assert v in adj[off[u]:off[u + 1]]
    g1 = qc.Graph.generate(1, 2)
    off1, _ = g1.csr()
    assert g1.n == 8000 and abs(int(off1[-1]) // 2 - 1_000_000) < 1000


def test_generator_config3_instances():
    seen = set()
    for i in range(4):
        g = qc.Graph.generate(3, 4, i)
        assert 500 <= g.n <= 20000 and 0.6 <= g.gamma <= 0.95 and 8 <= g.min_size <= 40
        seen.add(g.checksum())
    assert len(seen) == 4


def test_serialized_round_trip(tmp_path):
    off, adj = planted(60, 0.1, [8], 0.9, 3)
    g = qc.Graph.from_csr(off, adj)
    p = tmp_path / "g.txt"
    g.write_serialized(p)
    lines = p.read_text().split("\n")
    n, E, L2 = int(lines[0]), int(lines[1]), int(lines[2])
    assert n == 60 and E == len(adj)
    off2, adj2 = oracle.build_n2(off, adj)
    assert L2 == len(adj2)
    assert [int(x) for x in lines[5].split()] == [int(x) for x in adj2]
    h = qc.Graph.load_serialized(p)
    o2, a2 = h.csr()
    assert np.array_equal(o2, off) and np.array_equal(a2, adj)


def _csr_from_pairs(n, pairs):
    nbr = [set() for _ in range(n)]
    for u, v in pairs:
        if u != v:
            nbr[u].add(v)
            nbr[v].add(u)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in nbr])
    adj = np.array([x for s in nbr for x in sorted(s)], np.uint32)
    return off, adj


@pytest.mark.parametrize("case", ["hub_isolated", "n70", "sparse"])
def test_serialized_n2_records_equal_the_oracle(tmp_path, case):
    """Both N2 records of the serialized file (lists and offsets) against the
    oracle's N2 builder, itself pinned against a dense A + A^2.  The device
    writer is compared byte for byte with this file on the GPU."""
    rng = np.random.default_rng(11)
    if case == "hub_isolated":
        n = 300
        pairs = [(0, int(v)) for v in rng.choice(np.arange(1, 290), 150, replace=False)]
        pairs += [tuple(int(x) for x in rng.integers(0, 290, 2)) for _ in range(400)]
    elif case == "n70":
        n = 70
        pairs = [(i, (i * 7 + 3) % 70) for i in range(70)]
    else:
        n = 500
        pairs = [tuple(int(x) for x in rng.integers(0, n, 2)) for _ in range(300)]
    off, adj = _csr_from_pairs(n, pairs)
    g = qc.Graph.from_csr(off, adj)
    p = tmp_path / "g.txt"
    g.write_serialized(p)
    lines = p.read_text().split("\n")
    off2, adj2 = oracle.build_n2(off, adj)
    assert int(lines[2]) == len(adj2)
    assert [int(x) for x in lines[5].split()] == [int(x) for x in adj2]
    assert [int(x) for x in lines[6].split()] == [int(x) for x in off2]


def test_binary_round_trip(tmp_path):
    off, adj = planted(80, 0.1, [9], 0.9, 4)
    g = qc.Graph.from_csr(off, adj)
    p = tmp_path / "g.qcg"
    g.write_binary(p)
    h = qc.Graph.load_binary(p)
    o2, a2 = h.csr()
    assert np.array_equal(o2, off) and np.array_equal(a2, adj)
    assert h.checksum() == g.checksum()


def test_adjlist_load(tmp_path):
    p = tmp_path / "adj.txt"
    p.write_text("1 2\n0 2\n0 1 3\n2\n")
    g = qc.Graph.load_adjlist(p)
    off, adj = g.csr()
    assert list(off) == [0, 2, 4, 7, 8]
    assert list(adj) == [1, 2, 0, 2, 0, 1, 3, 2]


def test_corrupt_files(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("3\n4\n0\n1 0 2\n")
    with pytest.raises(qc.QCError) as e:
        qc.Graph.load_serialized(p)
    assert e.value.code == qc.QC_EIO
    with pytest.raises(qc.QCError) as e:
        qc.Graph.load_serialized(tmp_path / "missing.txt")
    assert e.value.code == qc.QC_EIO
    q = tmp_path / "bad.qcg"
    q.write_bytes(b"QCG1\x05\x00\x00\x00")
    with pytest.raises(qc.QCError):
        qc.Graph.load_binary(q)


@pytest.mark.parametrize("off,adj", [
    ([0, 1, 1], [1]),            # asymmetric
    ([0, 1, 2], [0, 0]),         # self loop
    ([0, 2, 3, 4], [2, 1, 0, 0]),  # unsorted list
    ([0, 1, 2], [5, 0]),         # out of range
])
def test_from_csr_validation(off, adj):
    with pytest.raises(qc.QCError) as e:
        qc.Graph.from_csr(np.array(off, np.uint64), np.array(adj, np.uint32))
    assert e.value.code == qc.QC_EINVAL


def test_parameter_validation_before_device():
    g = qc.Graph.generate(0, 1)
    for gamma, ms in [(0.4, 10), (1.1, 10), (0.9, 1), (0.9, 0)]:
        with pytest.raises(qc.QCError) as e:
            qc.mine(g, gamma, ms)
        assert e.value.code == qc.QC_EINVAL


@pytest.mark.skipif(qc.device_available(), reason="checks the no-device error path")
def test_mining_without_device_fails_loudly():
    g = qc.Graph.generate(0, 1)
    with pytest.raises(qc.QCError) as e:
        qc.mine(g, 0.9, 10)
    assert e.value.code == qc.QC_EDEVICE


@pytest.mark.skipif(qc.device_available(), reason="checks the no-device error path")
def test_device_n2_writer_without_device_fails_loudly(tmp_path):
    g = qc.Graph.generate(0, 1)
    with pytest.raises(qc.QCError) as e:
        g.write_serialized(tmp_path / "g.txt", device=True)
    assert e.value.code == qc.QC_EDEVICE
    assert not (tmp_path / "g.txt").exists()


def _cli(*args, cwd=None):
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=120)
    return r.returncode, r.stdout


def test_cli_usage_and_argument_errors(tmp_path):
    rc, out = _cli()
    assert rc == 1 and out.startswith("Usage: ./main <graph_file> <gamma> <min_size>")
    rc, out = _cli(tmp_path / "nope", 0.9, 10, tmp_path / "o.txt", 0)
    assert rc == 1 and out.strip() == "invalid graph file"
    g = qc.Graph.generate(0, 1)
    p = tmp_path / "g.txt"
    g.write_serialized(p)
    rc, out = _cli(p, 0.45, 10, tmp_path / "o.txt", 0)
    assert rc == 1 and "between .5 and 1" in out
    rc, out = _cli(p, 0.9, 1, tmp_path / "o.txt", 0)
    assert rc == 1 and "greater than 1" in out


@pytest.mark.skipif(qc.device_available(), reason="checks the no-device error path")
def test_cli_without_device(tmp_path):
    g = qc.Graph.generate(0, 1)
    p = tmp_path / "g.txt"
    g.write_serialized(p)
    rc, out = _cli(p, 0.9, 10, tmp_path / "o.txt", 3)
    assert rc == 1
    assert "scheduling toggle must be 0 or 1" in out  # warns only, like main.cu:451-453
    assert ">:PRE-PROCESSING" in out and "--->:LOADING TIME:" in out
    assert not (tmp_path / "o.txt").exists()
