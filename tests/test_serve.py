"""Scheduler of the tile-serving loop (CPU only, injected tile pipeline):
slot and HBM-budget admission, FIFO order, failure isolation, latency
statistics."""
import threading
import time

import pytest

from peakpatch_hip.serve import TileRequest, TileResult, TileServer, footprint_bytes, poisson_stream, replay, summarize


def fake_runner(service_s=0.05, fail_ids=()):
    live = {"now": 0, "max": 0}
    lock = threading.Lock()

    def run(state, req):
        with lock:
            live["now"] += 1
            live["max"] = max(live["max"], live["now"])
        try:
            time.sleep(service_s * req.nmesh / 128)
            if req.tile_id in fail_ids:
                raise ValueError("injected tile failure")
            return TileResult(req.tile_id, req.nmesh, npeaks=req.nmesh, nhalo=req.nmesh // 2)
        finally:
            with lock:
                live["now"] -= 1
    return run, live


def test_slots_bound_concurrency_and_keep_order():
    run, live = fake_runner()
    with TileServer(slots=3, run_tile=run) as srv:
        futs = [srv.submit(TileRequest(q, 128)) for q in range(9)]
        res = [f.result(timeout=30) for f in futs]
    assert [r.tile_id for r in res] == list(range(9))
    assert live["max"] == 3 and srv.max_in_flight == 3
    assert all(r.error is None and r.latency >= 0.05 for r in res)
    # FIFO admission: tiles start in arrival order
    starts = [r.start for r in res]
    assert starts == sorted(starts)


def test_hbm_budget_limits_admission():
    run, live = fake_runner()
    budget = 2 * footprint_bytes(128) + 1
    with TileServer(slots=4, run_tile=run, budget_bytes=budget) as srv:
        res = [f.result(timeout=30) for f in [srv.submit(TileRequest(q, 128)) for q in range(6)]]
    assert live["max"] == 2 and all(r.error is None for r in res)


def test_oversized_tile_fails_fast_and_others_run():
    run, _ = fake_runner()
    with TileServer(slots=2, run_tile=run, budget_bytes=footprint_bytes(256)) as srv:
        futs = [srv.submit(TileRequest(0, 128)), srv.submit(TileRequest(1, 768)), srv.submit(TileRequest(2, 256))]
        res = [f.result(timeout=30) for f in futs]
    assert res[1].error and res[1].error.startswith("AdmissionError")
    assert res[0].error is None and res[2].error is None


def test_failed_tile_is_isolated():
    run, _ = fake_runner(fail_ids=(1,))
    with TileServer(slots=2, run_tile=run) as srv:
        res = [f.result(timeout=30) for f in [srv.submit(TileRequest(q, 128)) for q in range(4)]]
    assert "injected tile failure" in res[1].error
    assert [r.error for r in (res[0], res[2], res[3])] == [None, None, None]


def test_replay_and_summary():
    run, _ = fake_runner(service_s=0.02)
    stream = poisson_stream(12, rate=50.0, seed=3, nmesh=(128, 256))
    assert all(b[0] >= a[0] for a, b in zip(stream, stream[1:]))
    with TileServer(slots=2, run_tile=run) as srv:
        res = replay(srv, stream)
    s = summarize(res)
    assert s["completed"] == 12 and s["failed"] == 0
    assert s["p99_ms"] >= s["p50_ms"] > 0 and s["tiles_per_s"] > 0
    assert s["peaks"] == sum(r.nmesh for r in res)


def test_submit_after_close_raises():
    run, _ = fake_runner()
    srv = TileServer(slots=1, run_tile=run)
    srv.close()
    with pytest.raises(RuntimeError):
        srv.submit(TileRequest(0, 128))


# ---- the native serving entry points through the C ABI (no GPU: device -1) ----
import ctypes as C  # noqa: E402

from peakpatch_hip import _lib as L  # noqa: E402


class _Native:
    """pp_serve_* with a ctypes tile function that sleeps `service_s` and
    records which requests ran."""

    def __init__(self, slots=2, queue=64, budget=1 << 40, service_s=0.05):
        self.lib = L.load()
        self.ran = []
        self.lock = threading.Lock()

        def fn(user, ctx, slot, ticket, request):
            assert ctx is None
            time.sleep(service_s)
            with self.lock:
                self.ran.append(int(request or 0))
            return 0
        self.fn = L.TILE_FN(fn)
        cfg = L.ServeConfig(-1, slots, queue, budget, 0.0, self.fn, None)
        self.h = C.c_void_p()
        L.check(self.lib.pp_serve_create(C.byref(cfg), C.byref(self.h)))

    def submit(self, rid, need=1):
        t = C.c_int64()
        rc = self.lib.pp_serve_submit(self.h, C.c_void_p(rid), need, C.byref(t))
        return rc, t.value

    def poll(self, t, timeout_ms=-1):
        st = C.c_int()
        times = (C.c_double * 4)()
        rc = self.lib.pp_serve_poll(self.h, t, timeout_ms, C.byref(st), times)
        return rc, st.value, list(times)

    def close(self):
        L.check(self.lib.pp_serve_destroy(self.h))


def test_native_injected_fault_leaves_other_tiles_intact():
    s = _Native(slots=2)
    try:
        tickets = [s.submit(100 + q)[1] for q in range(6)]
        L.check(s.lib.pp_serve_inject_fault(s.h, tickets[4]))
        out = [s.poll(t) for t in tickets]
        assert all(rc == 0 for rc, _, _ in out)
        assert [st for _, st, _ in out] == [0, 0, 0, 0, -9, 0]
        assert sorted(s.ran) == [100, 101, 102, 103, 105]
        buf = C.create_string_buffer(256)
        L.check(s.lib.pp_serve_message(s.h, tickets[4], buf, 256))
        assert b"fault injected" in buf.value
        stats = (C.c_int64 * 7)()
        L.check(s.lib.pp_serve_stats(s.h, stats))
        assert stats[3] == 6 and stats[4] == 1 and stats[2] == 2
    finally:
        s.close()


def test_native_busy_timeout_and_capacity():
    s = _Native(slots=1, queue=2, budget=1000, service_s=0.3)
    try:
        rc0, t0 = s.submit(1)          # starts at once
        time.sleep(0.05)
        assert s.submit(2)[0] == 0 and s.submit(3)[0] == 0
        assert s.submit(4)[0] == -7    # PP_ERR_BUSY: two already wait
        assert s.submit(5, need=1001)[0] == -6   # PP_ERR_CAPACITY: above the budget
        rc, st, _ = s.poll(t0, timeout_ms=10)
        assert rc == -8                # PP_ERR_TIMEOUT: still running
        rc, st, times = s.poll(t0)
        assert rc == 0 and st == 0 and times[2] >= times[1] >= times[0]
    finally:
        s.close()
    assert s.ran == [1, 2, 3]


def test_tileserver_inject_fault():
    run, _ = fake_runner(service_s=0.05)
    with TileServer(slots=1, run_tile=run) as srv:
        futs = [srv.submit(TileRequest(q, 128)) for q in range(4)]
        srv.inject_fault(futs[3])
        res = [f.result(timeout=30) for f in futs]
    assert [r.error is None for r in res] == [True, True, True, False]
    assert res[3].error.startswith("PP_ERR_INJECTED")
