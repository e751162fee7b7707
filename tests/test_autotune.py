"""Autotuner search logic on the CPU with a stand-in timing function (the
GPU path is exercised by the GPU tests and bench.py)."""
import cuqc_amd.autotune as at


class _Res:
    def __init__(self, t):
        self.stats = {k: 1 for k in at._INVARIANT}
        self.stats["t_kernel_ms"] = t
        self.raw = [[1, 2, 3]]


def test_coordinate_descent_revisits_parameters(monkeypatch):
    # the best point needs split_chunk = -1 together with lds_wide = 0: a
    # single sweep that fixes lds_wide first (under the default chunk) misses it
    def fake_mine(g, ga, ms, sched, mode, window, **kw):
        t = 100.0
        if kw["split_chunk"] == -1:
            t = 50.0 if kw["lds_wide"] == 0 else 80.0
        elif kw["lds_wide"] == 1024:
            t = 90.0
        return _Res(t)

    import cuqc_amd
    monkeypatch.setattr(cuqc_amd, "mine", fake_mine)

    class G:
        def checksum(self):
            return 1

    s = at.tune([G()], [0.9], [10], reps=1)
    assert s.split_chunk == -1 and s.lds_wide in (0, 256, 512)
    assert s.kernel_ms == 50.0


def test_schedule_changing_the_result_raises(monkeypatch):
    calls = {"n": 0}

    def fake_mine(g, ga, ms, sched, mode, window, **kw):
        calls["n"] += 1
        r = _Res(10.0)
        if calls["n"] > 1:
            r.stats["tasks"] = 2
        return r

    import cuqc_amd
    monkeypatch.setattr(cuqc_amd, "mine", fake_mine)

    class G:
        def checksum(self):
            return 1

    try:
        at.tune([G()], [0.9], [10], reps=1)
    except RuntimeError as e:
        assert "changed the mining result" in str(e)
    else:
        raise AssertionError("expected RuntimeError")


def test_lucky_single_timing_does_not_leave_the_defaults(monkeypatch):
    # split_chunk 128 times fast once, then slower than the defaults: the
    # confirmation runs keep the defaults
    seen = {"n": 0}

    def fake_mine(g, ga, ms, sched, mode, window, **kw):
        if kw["split_chunk"] == 128:
            seen["n"] += 1
            return _Res(90.0 if seen["n"] == 1 else 110.0)
        return _Res(100.0)

    import cuqc_amd
    monkeypatch.setattr(cuqc_amd, "mine", fake_mine)

    class G:
        def checksum(self):
            return 1

    s = at.tune([G()], [0.9], [10], reps=1)
    assert s.split_chunk == at.DEFAULT["split_chunk"]
    assert s.kernel_ms == 100.0
