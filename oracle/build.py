"""Build the oracle's C++ restatement into oracle/_build/liboracle.so.

Test infrastructure only (see qc_oracle.cpp).  The reference's own sources
cannot be built here: main.cu needs nvcc and the CUDA runtime, and
adjToSer.cpp includes cuda_runtime.h, so there is no oracle/_ref build.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build" / "liboracle.so"


def build(force: bool = False) -> Path:
    src = HERE / "qc_oracle.cpp"
    OUT.parent.mkdir(exist_ok=True)
    if force or not OUT.exists() or OUT.stat().st_mtime < src.stat().st_mtime:
        tmp = str(OUT) + ".tmp"
        subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", str(src), "-o", tmp], check=True)
        Path(tmp).replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build("--force" in sys.argv))
