"""qc_remove_nonmax_noalloc (RemoveNonMax with no device allocation) against
the allocating entry point qc_remove_nonmax, through the C ABI.

The caller's device buffers come from hipMalloc of the HIP runtime the
library itself is linked against (a second runtime in the process, such as
the one torch bundles, cannot share the device); the entry point uses only
the library's fixed scratch array.  Two calls in a row, the second on a
smaller input, must each match qc_remove_nonmax exactly (same sets, same
lexicographic order): the scratch keeps the first call's contents, so any
state that is not cleared would show up in the second call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import cuqc_amd as qc

pytestmark = pytest.mark.gpu

QC_OK, QC_EINVAL, QC_EOVERFLOW = 0, 1, 4


def _fn():
    f = qc.load_library().qc_remove_nonmax_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p]
    return f


def _raw_sets(seed: int, k: int, n: int) -> list[np.ndarray]:
    """Raw sets with the shapes the miner emits: duplicates (in another vertex
    order), proper subsets and unrelated sets."""
    rng = np.random.default_rng(seed)
    sets: list[np.ndarray] = []
    for _ in range(k):
        r = rng.random()
        if sets and r < 0.25:
            sets.append(rng.permutation(sets[int(rng.integers(len(sets)))]))
        elif sets and r < 0.5:
            base = sets[int(rng.integers(len(sets)))]
            if len(base) > 2:
                drop = int(rng.integers(1, len(base) - 1))
                sets.append(rng.permutation(base)[drop:])
                continue
            sets.append(rng.choice(n, size=3, replace=False))
        else:
            sets.append(rng.choice(n, size=int(rng.integers(2, 70)), replace=False))
    return [np.asarray(s, np.uint32) for s in sets]


_HIP = None


def _hip() -> C.CDLL:
    """The HIP runtime libqcmine.so is linked against (found among the
    process's mapped libraries once the library is loaded)."""
    global _HIP
    if _HIP is None:
        qc.load_library()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        assert path, "libamdhip64 is not loaded"
        H = C.CDLL(path)
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP = H
    return _HIP


class _Dev:
    """A device buffer holding `arr` (copied in), read back with get()."""

    def __init__(self, arr: np.ndarray):
        self.arr = np.ascontiguousarray(arr)
        self.p = C.c_void_p()
        H = _hip()
        assert H.hipMalloc(C.byref(self.p), max(self.arr.nbytes, 16)) == 0
        assert H.hipMemcpy(self.p, self.arr.ctypes.data, self.arr.nbytes, 1) == 0  # host to device

    def get(self) -> np.ndarray:
        out = np.empty_like(self.arr)
        assert _hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        _hip().hipFree(self.p)


def _run_noalloc(sets, n_vertices: int):
    k = len(sets)
    off = np.zeros(k + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in sets])
    nv = int(off[-1])
    v = np.concatenate(sets).astype(np.uint32) if k else np.zeros(1, np.uint32)
    # outputs prefilled with garbage: every entry that is read must be written
    bufs = [_Dev(off), _Dev(v), _Dev(np.full(1, 12345, np.uint64)), _Dev(np.full(k + 1, 7, np.uint64)),
            _Dev(np.full(max(nv, 1), 7, np.uint32))]
    d_off, d_v, d_cnt, d_ooff, d_ov = bufs
    try:
        rc = _fn()(k, nv, d_off.p, d_v.p, n_vertices, d_cnt.p, d_ooff.p, d_ov.p, None)
        assert _hip().hipDeviceSynchronize() == 0
        if rc != QC_OK:
            return rc, None
        cnt = int(d_cnt.get()[0])
        if cnt == 2 ** 64 - 1:
            return rc, -1
        oo = d_ooff.get().astype(np.int64)
        ov = d_ov.get().astype(np.int64)
        return rc, [tuple(int(x) for x in ov[oo[i]:oo[i + 1]]) for i in range(cnt)]
    finally:
        for b in bufs:
            b.free()


def _expected(sets):
    return [tuple(int(x) for x in s) for s in qc.remove_nonmax(sets)]


def test_noalloc_matches_remove_nonmax_twice():
    if not qc.device_available():
        pytest.skip("no gfx950 device")
    big = _raw_sets(11, 3000, 5000)
    small = _raw_sets(12, 400, 900)
    for name, sets, n in (("first call, 3000 sets", big, 5000), ("second call, 400 sets", small, 900)):
        print(f"[no_alloc] {name}", flush=True)
        rc, got = _run_noalloc(sets, n)
        assert rc == QC_OK, qc.load_library().qc_last_error()
        exp = _expected(sets)
        assert len(exp) > 0
        assert got == exp, name


def test_noalloc_miner_raw_sets():
    """Raw sets of a real mining run (workload 0) as the input."""
    if not qc.device_available():
        pytest.skip("no gfx950 device")
    g = qc.Graph.generate(0, 1)
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="persistent")
    raw = [np.asarray(s, np.uint32) for s in res.raw]
    rc, got = _run_noalloc(raw, 200)
    assert rc == QC_OK, qc.load_library().qc_last_error()
    assert got == _expected(raw)
    assert got == [tuple(int(x) for x in s) for s in res.maximal]


def test_noalloc_limits():
    """Out-of-range ids give count UINT64_MAX (no out-of-bounds access);
    inputs beyond the scratch size are rejected before any launch; an empty
    input gives zero sets."""
    if not qc.device_available():
        pytest.skip("no gfx950 device")
    sets = _raw_sets(13, 50, 100)
    rc, cnt = _run_noalloc(sets, 60)
    assert rc == QC_OK and cnt == -1
    rc, _ = _run_noalloc(sets, 1 << 31)
    assert rc == QC_EOVERFLOW
    assert "limit" in qc.load_library().qc_last_error().decode()
    rc, got = _run_noalloc([], 10)
    assert rc == QC_OK and got == []
