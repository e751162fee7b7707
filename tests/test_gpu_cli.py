"""The drop-in boundary end to end on the GPU: the qcmine command line
(main.cu:416-519) on serialized text (main.cu:96-122) and binary graph files,
in both scheduling toggles and both execution forms.

The output file must equal, byte for byte, the file the oracle's maximal
sets give in the reference's OutputOneSet format (main.cu:4641-4651); the raw
dump (QCMINE_RAW, the temp.txt format of flush_cliques, main.cu:1234-1246)
must hold the oracle's raw multiset; stdout must carry the reference's
markers in order; QCMINE_DEBUG=1 adds the size report (main.cu:2635-2646)
whose per-level task counts add up to the oracle's.
"""
from __future__ import annotations

import os
import re
import subprocess
import sys

import pytest

import cuqc_amd as qc
import oracle
from cuqc_amd import workloads

pytestmark = pytest.mark.gpu

CLI = str(qc.library_path().parent / "qcmine")
MARKERS = [">:PRE-PROCESSING", "--->:LOADING TIME:", ">:INITIALIZING TASKS", ">:BEGINNING EXPANSION",
           "--->:ENUMERATION TIME:", ">:REMOVING NON-MAXIMAL CLIQUES", ">:NUMBER OF MAXIMAL CLIQUES:",
           "--->:REMOVE NON-MAX TIME:", "--->:TOTAL TIME:", ">:PROGRAM END"]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _expected_file(sets) -> bytes:
    return "".join(f"{len(s)} " + "".join(f"{int(x)} " for x in s) + "\n" for s in sets).encode()


def _raw_sets(path) -> list[tuple[int, ...]]:
    out = []
    with open(path) as f:
        for line in f:
            x = [int(t) for t in line.split()]
            assert len(x) == x[0] + 1, line  # "len v1 v2 ... " per line
            out.append(tuple(sorted(x[1:])))
    return sorted(out)


def _graphs():
    g0 = qc.Graph.generate(0, 1)
    g3 = workloads.graph(workloads.WORKLOADS[3], 12)  # batch instance 12: 3,649 tasks at full depth
    return [("config0", g0), ("batch_instance12", g3)]


@pytest.fixture(scope="module")
def cases(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    out = []
    for name, g in _graphs():
        off, adj = g.csr()
        ref = oracle.mine(off, adj, g.gamma, g.min_size)
        ser, binf = d / f"{name}.txt", d / f"{name}.qcg"
        g.write_serialized(ser)
        g.write_binary(binf)
        out.append((name, g, ref, ser, binf))
    return out


@pytest.mark.parametrize("mode", ["persistent", "plain"])
@pytest.mark.parametrize("toggle", [0, 1])
def test_cli_end_to_end(cases, tmp_path, mode, toggle):
    for name, g, ref, ser, binf in cases:
        for fmt, path in (("text", ser), ("binary", binf)):
            _say(f"[gpu] qcmine {name} ({fmt}) gamma {g.gamma} min_size {g.min_size} toggle {toggle} {mode}")
            outp, rawp = tmp_path / f"{name}_{fmt}.out", tmp_path / f"{name}_{fmt}.raw"
            env = dict(os.environ, QCMINE_MODE=mode, QCMINE_RAW=str(rawp), QCMINE_DEBUG="1")
            r = subprocess.run([CLI, str(path), repr(g.gamma), str(g.min_size), str(outp), str(toggle)],
                               capture_output=True, text=True, timeout=120, env=env)
            assert r.returncode == 0, r.stdout + r.stderr
            # markers in the reference's order
            at = 0
            for m in MARKERS:
                k = r.stdout.find(m, at)
                assert k >= 0, (m, r.stdout)
                at = k
            mx = oracle.nonmax(ref["raw"])
            assert f">:NUMBER OF MAXIMAL CLIQUES: {len(mx)}" in r.stdout
            assert outp.read_bytes() == _expected_file(mx), (name, fmt)
            assert _raw_sets(rawp) == oracle.canon(ref["raw"]), (name, fmt)
            # size report: T1 counts over the levels add up to the tasks
            levels = re.findall(r"^L: (\d+) T1: (\d+) (\d+) T2: (\d+) (\d+) B: (\d+) (\d+) C: (\d+) (\d+)$",
                                r.stdout, flags=re.M)
            assert levels, r.stdout
            assert sum(int(x[1]) for x in levels) == ref["stats"]["tasks"]
            assert sum(int(x[2]) for x in levels) == ref["stats"]["sum_t"]
            assert sum(int(x[3]) for x in levels) == ref["stats"]["children"]
            assert "TASKS SIZE:" in r.stdout and "CLIQUES SIZE:" in r.stdout


def test_cli_empty_result(tmp_path):
    """No quasi-clique: the reference's error message and return code 1, no
    output file (main.cu:4833, README.md:108)."""
    g = qc.Graph.generate(0, 1)
    p = tmp_path / "g.txt"
    g.write_serialized(p)
    _say("[gpu] qcmine config0 with min_size 60 (no quasi-clique)")
    r = subprocess.run([CLI, str(p), "0.9", "60", str(tmp_path / "o.txt"), "0"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1
    assert "Error: no quasi-cliques found" in r.stdout
    assert not (tmp_path / "o.txt").exists()
