"""Dataset converters (the reference's edgeToAdj.py, rmEmptyVert.py and
adjToSer.cpp): hand-worked known answers, the scripts' behaviour restated in
a few lines of Python on seeded random inputs, and the serialized output
loaded back (CPU only)."""
import os
import random
import subprocess

import numpy as np
import pytest

import cuqc_amd as qc
import oracle

CONV = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cuqc_amd", "qcconv")


def _conv(tmp_path, fn, text, *args):
    src = tmp_path / "in.txt"
    src.write_bytes(text.encode() if isinstance(text, str) else text)
    dst = tmp_path / "out.txt"
    fn(src, dst, *args)
    return dst.read_text()


# ---- edgeToAdj.py ----------------------------------------------------------

def test_edge_to_adj_known_answers(tmp_path):
    edges = "0 1\n1 2\n0 2\n"
    assert _conv(tmp_path, qc.convert_edge_to_adj, edges, False) == "1 2\n2\n\n"
    assert _conv(tmp_path, qc.convert_edge_to_adj, edges, True) == "1 2\n0 2\n1 0\n"
    # ids are kept as given; vertices without edges print empty lines
    assert _conv(tmp_path, qc.convert_edge_to_adj, "3 0", False) == "\n\n\n0\n"
    # an empty edge list prints vertex 0's empty line
    assert _conv(tmp_path, qc.convert_edge_to_adj, "", False) == "\n"


def _edge_to_adj_model(lines, duplicate):
    adj, mx = {}, 0
    for u, v in lines:
        mx = max(mx, u, v)
        adj.setdefault(u, []).append(v)
        if duplicate:
            adj.setdefault(v, []).append(u)
    return "".join(" ".join(map(str, adj.get(x, []))) + "\n" for x in range(mx + 1))


@pytest.mark.parametrize("seed", range(6))
def test_edge_to_adj_random(tmp_path, seed):
    rng = random.Random(seed)
    pairs = [(rng.randrange(40), rng.randrange(40)) for _ in range(rng.randrange(1, 120))]
    text = "".join(f"{u}{rng.choice([' ', '  ', chr(9)])}{v}{rng.choice(['', ' '])}\n" for u, v in pairs)
    for dup in (False, True):
        assert _conv(tmp_path, qc.convert_edge_to_adj, text, dup) == _edge_to_adj_model(pairs, dup)


def test_edge_to_adj_rejects_malformed_lines(tmp_path):
    for bad in ["0 1 2\n", "0\n", "0 x\n", "0 1\n\n1 2\n"]:
        with pytest.raises(qc.QCError):
            _conv(tmp_path, qc.convert_edge_to_adj, bad, False)


# ---- rmEmptyVert.py --------------------------------------------------------

def test_rm_empty_known_answer(tmp_path):
    assert _conv(tmp_path, qc.convert_rm_empty, "2\n\n0\n") == "1\n0\n"
    assert _conv(tmp_path, qc.convert_rm_empty, "\n\n") == ""
    with pytest.raises(qc.QCError):  # 1 names an empty line (the script's KeyError)
        _conv(tmp_path, qc.convert_rm_empty, "1\n\n0\n")


@pytest.mark.parametrize("seed", range(4))
def test_rm_empty_random(tmp_path, seed):
    rng = random.Random(100 + seed)
    n = 30
    keep = [i for i in range(n) if rng.random() < 0.6]
    lists = [[rng.choice(keep) for _ in range(rng.randrange(1, 6))] if i in keep else [] for i in range(n)]
    text = "".join(" ".join(map(str, l)) + "\n" for l in lists)
    vmap = {v: k for k, v in enumerate(keep)}
    want = "".join(" ".join(str(vmap[x]) for x in l) + "\n" for l in lists if l)
    assert _conv(tmp_path, qc.convert_rm_empty, text) == want


# ---- adjToSer.cpp ----------------------------------------------------------

def test_adj_to_ser_known_answers(tmp_path):
    tri = "1 2\n0 2\n0 1"
    assert _conv(tmp_path, qc.convert_adj_to_ser, tri) == "3\n6\n6\n1 2 0 2 0 1\n0 2 4 6\n1 2 0 2 0 1\n0 2 4 6\n"
    # a final line break closes one more (empty) vertex, as adjToSer.cpp:105-110 does
    assert _conv(tmp_path, qc.convert_adj_to_ser, tri + "\n") == \
        "4\n6\n6\n1 2 0 2 0 1\n0 2 4 6 6\n1 2 0 2 0 1\n0 2 4 6 6\n"
    # a path 0-1-2: N2(0) = {1, 2}
    assert _conv(tmp_path, qc.convert_adj_to_ser, "1\n0 2\n1") == "3\n4\n6\n1 0 2 1\n0 1 3 4\n1 2 0 2 0 1\n0 2 4 6\n"
    # a space always closes a number, so "1 " then the line end adds a 0 (a
    # self loop, kept in the 1-hop part of N2(0))
    assert _conv(tmp_path, qc.convert_adj_to_ser, "1 \n0") == "2\n3\n3\n0 1 0\n0 2 3\n0 1 0\n0 2 3\n"


def _clean_adj_text(rng, n, p):
    adj = [set() for _ in range(n)]
    for u in range(n):
        for v in range(u + 1, n):
            if rng.random() < p:
                adj[u].add(v)
                adj[v].add(u)
    lines = []
    for u in range(n):
        l = list(adj[u])
        rng.shuffle(l)
        lines.append(" ".join(map(str, l)))
    return "\n".join(lines), adj


@pytest.mark.parametrize("seed", range(5))
def test_adj_to_ser_matches_n2_rule_and_loads(tmp_path, seed):
    rng = random.Random(seed)
    n = 40
    text, adj = _clean_adj_text(rng, n, 0.1)
    # no isolated vertices at the end keeps the file free of parse quirks
    out = _conv(tmp_path, qc.convert_adj_to_ser, text)
    rec = out.split("\n")
    assert int(rec[0]) == n
    off = np.array([0] + list(np.cumsum([len(a) for a in adj])), np.uint64)
    nbrs = np.array([x for a in adj for x in sorted(a)], np.uint32)
    assert [int(x) for x in rec[3].split()] == [int(x) for x in nbrs]
    off2, adj2 = oracle.build_n2(off, nbrs)
    assert [int(x) for x in rec[5].split()] == [int(x) for x in adj2]
    assert [int(x) for x in rec[6].split()] == [int(x) for x in off2]
    g = qc.Graph.load_serialized(tmp_path / "out.txt")
    o2, a2 = g.csr()
    assert np.array_equal(o2, off) and np.array_equal(a2, nbrs)


def test_qcconv_cli(tmp_path):
    src = tmp_path / "e.txt"
    src.write_text("0 1\n1 2\n")
    r = subprocess.run([CONV, "edge-to-adj", str(src), "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout == "1\n0 2\n1\n"
    adj = tmp_path / "a.txt"
    adj.write_text(r.stdout.rstrip("\n"))
    ser = tmp_path / "s.txt"
    r = subprocess.run([CONV, "adj-to-ser", str(adj), str(ser)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and ser.read_text().startswith("3\n4\n")
    r = subprocess.run([CONV, "adj-to-ser", str(tmp_path / "missing"), str(ser)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and r.stdout.strip() == "invalid graph file"
    r = subprocess.run([CONV], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stdout


# ---- against the reference's own scripts ------------------------------------

def _fixtures():
    import json
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "convert_fixtures.json")
    with open(path) as f:
        return json.load(f)


@pytest.mark.parametrize("i", range(18))
def test_edge_to_adj_matches_reference_script(tmp_path, i):
    """Bytes printed by edgeToAdj.py itself (tests/golden/make_convert_fixtures.py)."""
    case = _fixtures()["edge_to_adj"][i]
    assert _conv(tmp_path, qc.convert_edge_to_adj, case["input"], bool(case["duplicate"])) == case["output"]


@pytest.mark.parametrize("i", range(6))
def test_rm_empty_matches_reference_script(tmp_path, i):
    """Bytes written by rmEmptyVert.py itself (tests/golden/make_convert_fixtures.py)."""
    case = _fixtures()["rm_empty"][i]
    assert _conv(tmp_path, qc.convert_rm_empty, case["input"]) == case["output"]
