"""Build libqcmine.so (gfx950 HIP kernels + C ABI) and the qcmine CLI in-tree.

    python -m cuqc_amd.build            # compile what changed
    python -m cuqc_amd.build --force    # recompile everything

Objects go to cuqc_amd/_build/, the library to cuqc_amd/libqcmine.so, the
CLI to cuqc_amd/qcmine and the dataset converter CLI to cuqc_amd/qcconv
(rpath $ORIGIN).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libqcmine.so"
CLI = PKG / "qcmine"
CONV = PKG / "qcconv"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["expand.hip", "root.hip", "driver.hip", "nonmax.hip", "collect.hip", "persist.hip", "mlp.hip",
               "n2.hip", "capi.hip"]
CXX_SOURCES = ["graph_io.cpp", "gen.cpp", "convert.cpp"]
HEADERS = ["qc_host.h", "qc_device.h", "qc_internal.h"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # pool blocks hold a free-list link in the bytes of a task record
    "-fno-strict-aliasing",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall"]


def _stale(obj: Path, src: Path) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    deps = [src, PKG.parent / "include" / "qcmine.h"] + [SRC / h for h in HEADERS]
    return any(d.exists() and d.stat().st_mtime > t for d in deps)


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{msg}")


def build(force: bool = False, jobs: int | None = None, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    tasks = []
    for s in HIP_SOURCES:
        src, obj = SRC / s, BUILD / (s + ".o")
        if force or _stale(obj, src):
            tasks.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for s in CXX_SOURCES:
        src, obj = SRC / s, BUILD / (s + ".o")
        if force or _stale(obj, src):
            tasks.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for cmd in tasks:
            if verbose:
                print(" ".join(cmd), flush=True)
        list(ex.map(_run, tasks))
    objs = [str(BUILD / (s + ".o")) for s in HIP_SOURCES + CXX_SOURCES]
    if force or tasks or not LIB.exists():
        tmp = str(LIB) + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
        os.replace(tmp, LIB)
    main_src = SRC / "qcmine_main.cpp"
    if force or not CLI.exists() or CLI.stat().st_mtime < max(LIB.stat().st_mtime, main_src.stat().st_mtime):
        _run(["g++", *CXX_FLAGS, str(main_src), "-o", str(CLI), f"-L{PKG}", "-lqcmine",
              "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib"])
    conv_src = SRC / "qcconv_main.cpp"
    if force or not CONV.exists() or CONV.stat().st_mtime < max(LIB.stat().st_mtime, conv_src.stat().st_mtime):
        _run(["g++", *CXX_FLAGS, str(conv_src), "-o", str(CONV), f"-L{PKG}", "-lqcmine",
              "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib"])
    return LIB


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    build(a.force, a.j, a.v)
    print(f"built {LIB}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
