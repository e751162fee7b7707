"""Write tests/golden/convert_fixtures.json: inputs and the outputs of the
reference's own preprocessing scripts, edgeToAdj.py and rmEmptyVert.py,
run as they are from /root/reference (only in the build container; the
fixture is data: each input next to the bytes the script printed or wrote).

    python tests/golden/make_convert_fixtures.py

The inputs are seeded: edge lists with repeated and reversed edges, self
loops, vertex gaps and mixed blank separators (in both duplicate modes), and
adjacency lists with empty lines between and after non-empty ones.
"""
from __future__ import annotations

import json
import os
import random
import subprocess
import sys
import tempfile

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def edge_inputs() -> list[str]:
    out = ["0 1\n1 2\n0 2\n", "3 0\n", "5 5\n2 7\n7 2\n2 7\n"]
    for seed in range(6):
        rng = random.Random(seed)
        n = rng.choice([10, 40, 200])
        pairs = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randrange(1, 150))]
        out.append("".join(f"{u}{rng.choice([' ', '  ', chr(9)])}{v}{rng.choice(['', ' '])}\n" for u, v in pairs))
    return out


def adj_inputs() -> list[str]:
    out = ["2\n\n0\n", "1 2\n0\n0\n\n\n"]
    for seed in range(4):
        rng = random.Random(100 + seed)
        n = rng.choice([12, 30, 80])
        keep = [i for i in range(n) if rng.random() < 0.6]
        lists = [[rng.choice(keep) for _ in range(rng.randrange(1, 6))] if i in keep else [] for i in range(n)]
        out.append("".join(" ".join(map(str, l)) + "\n" for l in lists))
    return out


def main() -> int:
    py = sys.executable
    cases = {"edge_to_adj": [], "rm_empty": []}
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "in.txt")
        for text in edge_inputs():
            with open(src, "w") as f:
                f.write(text)
            for dup in (0, 1):
                r = subprocess.run([py, os.path.join(REF, "edgeToAdj.py"), src, str(dup)], capture_output=True,
                                   text=True, check=True)
                cases["edge_to_adj"].append({"input": text, "duplicate": dup, "output": r.stdout})
        dst = os.path.join(d, "out.txt")
        for text in adj_inputs():
            with open(src, "w") as f:
                f.write(text)
            subprocess.run([py, os.path.join(REF, "rmEmptyVert.py"), src, dst], capture_output=True, check=True)
            with open(dst) as f:
                cases["rm_empty"].append({"input": text, "output": f.read()})
    with open(os.path.join(HERE, "convert_fixtures.json"), "w") as f:
        json.dump(cases, f, indent=1)
    print({k: len(v) for k, v in cases.items()})
    return 0


if __name__ == "__main__":
    sys.exit(main())
