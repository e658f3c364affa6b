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
