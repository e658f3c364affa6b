"""Tile-serving loop: a stream of independent light-cone tiles of uneven
nmesh, the unit of work of run_multitile_split (src/MultiResolution.jl:
719-1180, one `process_tile!` per tile; round-robin dispatch :1183-1203).

Design
------
* Slots.  Each slot owns a Context (its own HIP stream, pooled device
  buffers, the collapse table and the shell tables, all resident for the
  server's life) and one worker thread.  Everything a tile does -- input
  synthesis, filter bank + stencil, shell analysis, record packing -- is
  queued on that one stream, so tiles on
  different slots overlap on the device while each tile stays in order.
* Admission by free HBM.  A tile's device footprint is estimated from its
  nmesh (``footprint_bytes``).  The scheduler admits the oldest waiting tile
  when a slot is idle and the footprints of the tiles in flight plus this one
  fit the budget (a fraction of the HBM free when the server started).  A
  tile that could never fit is failed at once instead of blocking the queue.
* Continuous batching.  Requests arriving while others run are admitted as
  soon as a slot and memory free up; the device interleaves their kernels.
  Results do not depend on what else runs: every kernel of the path is
  deterministic and a tile touches only its own slot's buffers.
* Latency = completion - arrival (queueing included); the stream summary
  reports tiles/s and p50/p99.

The tile pipeline is injectable (``run_tile``) so the scheduler's policies
and failure handling are testable without a GPU.
"""
from __future__ import annotations

import collections
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field, replace

import numpy as np

from .pipeline import TileConfig

# c5 tiles: the c2 physics at fixed resolution (alatt = 250/256 Mpc/h), so
# every tile shares one shell table whatever its nmesh; light-cone redshifts
# from an observer at the origin.
SERVE_TEMPLATE = TileConfig("c5", 256, 250.0, 20, 1.0, 15.0, ilpt=2, ioutshear=1, ievol=1, z_max=4.6,
                            cen=(0.0, 0.0, 0.0))
SERVE_NMESH = (128, 256, 384, 512, 768)
# run_multitile_split runs behind the stream: one global field per nmesh,
# SPLIT_NTILE^3 tiles each, coarse grid COARSE_FACTOR x SPLIT_NTILE
SPLIT_NTILE = 4
COARSE_FACTOR = 5
SPLIT_SEED = 42


@dataclass
class TileRequest:
    tile_id: int
    nmesh: int
    center: tuple = (0.0, 0.0, 0.0)  # periodic tiles: the tile centre
    seed: int = 42                   # periodic tiles: the tile's own field
    arrival: float = 0.0             # perf_counter() at submission (set by submit)
    tile: tuple | None = None        # split tiles: (it, jt, kt), 1-based, in the nmesh's global grid


@dataclass
class TileResult:
    tile_id: int
    nmesh: int
    records: np.ndarray | None = None   # (nhalo, 33) float32 ExtHaloRecord
    RTHLmax: float = 0.0
    npeaks: int = 0
    nhalo: int = 0
    arrival: float = 0.0
    start: float = 0.0
    done: float = 0.0
    slot: int = -1
    error: str | None = None
    stages: dict = field(default_factory=dict)  # seconds per pipeline stage (host clock)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def tile_config(req: TileRequest, template: TileConfig = SERVE_TEMPLATE) -> TileConfig:
    alatt = template.alatt
    return replace(template, name=f"{template.name}/tile{req.tile_id}", n=req.nmesh, boxsize=req.nmesh * alatt,
                   tile_center=tuple(req.center), seed=req.seed)


def footprint_bytes(n: int) -> int:
    """Device bytes one tile holds at its peak: the 9 Float32 fields, their
    spectrum and the synthesis temporaries (~24 B/cell), filter scratch and
    mask (~9 B/cell), the transposed delta, profiles and results
    (~30 B/cell at ~1.5e-2 peaks/cell), plus fixed tables and plans."""
    return int(n) ** 3 * 100 + (256 << 20)


def poisson_stream(ntiles: int, rate: float, seed: int = 0, nmesh=SERVE_NMESH, weights=None,
                   extent: float = 1500.0, ntile: int = SPLIT_NTILE):
    """(arrival offset s, TileRequest) pairs: exponential gaps at `rate`
    tiles/s, nmesh drawn from `nmesh`, each request a uniformly drawn tile
    (it, jt, kt) of that nmesh's split run (and, for periodic tiles, a centre
    uniform in a light-cone cube of side `extent` Mpc/h)."""
    rng = np.random.default_rng(seed)
    w = None if weights is None else np.asarray(weights, float) / np.sum(weights)
    t = 0.0
    out = []
    for q in range(ntiles):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        n = int(rng.choice(nmesh, p=w))
        c = tuple(float(v) for v in rng.uniform(0.0, extent, 3))
        tl = tuple(int(v) for v in rng.integers(1, ntile + 1, 3))
        out.append((t, TileRequest(q, n, c, seed=1000 + q, tile=tl)))
    return out


class TileServer:
    """Continuous-batching tile server over the library's serving loop
    (pp_serve_*): scheduling, admission and the slot threads are native;
    ``run_tile(slot_state, request)`` returns a TileResult.

    Without ``run_tile`` the server runs the GPU pipeline of
    ``GpuTileRunner`` on the contexts the native server creates for its slots
    on ``device``, and admits by device memory: the budget is
    ``hbm_fraction`` of the free memory when the server starts (or
    ``budget_bytes``), and a tile also needs the live free memory to cover
    its footprint.  With ``run_tile`` the slots have no native context (the
    caller's ``slot_states`` carry whatever the pipeline needs) and the
    budget is ``budget_bytes`` (unbounded if None)."""

    def __init__(self, slots: int = 4, budget_bytes: int | None = None, run_tile=None, slot_states=None,
                 footprint=footprint_bytes, device: int = 0, hbm_fraction: float = 0.8, runner=None,
                 queue_capacity: int = 1 << 20):
        import ctypes as C

        from . import _lib as L
        self._L = L
        self._lib = L.load()
        self._runner = None
        self._own_runner = False
        native_ctx = run_tile is None
        if native_ctx:
            self._own_runner = runner is None
            self._runner = runner if runner is not None else GpuTileRunner(device)
            run_tile = self._runner.run
            if footprint is footprint_bytes:
                footprint = self._runner.footprint
        self._footprint = footprint
        if not native_ctx and slot_states is None:
            slot_states = [None] * slots
        self._run = run_tile
        self._lock = threading.Lock()
        self._reqs = {}
        self._next_id = 0
        self._closed = False
        self._unresolved = 0
        self._cb = L.TILE_FN(self._call)
        cfg = L.ServeConfig(device if native_ctx else -1, slots, queue_capacity,
                            0 if (native_ctx and budget_bytes is None) else int(budget_bytes or (1 << 62)),
                            float(hbm_fraction), self._cb, None)
        h = C.c_void_p()
        L.check(self._lib.pp_serve_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._t0 = time.perf_counter()
        if native_ctx:
            slot_states = []
            for q in range(slots):
                c = C.c_void_p()
                L.check(self._lib.pp_serve_context(self._h, q, C.byref(c)))
                slot_states.append(self._runner.make_slot(handle=c.value))
        self._states = slot_states
        self.budget = self.stats()["budget"]
        self._done_thread = threading.Thread(target=self._collect, name="tile-completions", daemon=True)
        self._done_thread.start()

    # ---------------------------------------------------------------- API
    def submit(self, req: TileRequest) -> Future:
        import ctypes as C
        fut: Future = Future()
        req.arrival = time.perf_counter()
        with self._lock:
            if self._closed:
                raise RuntimeError("server closed")
            rid = self._next_id
            self._next_id += 1
            self._reqs[rid] = (req, fut)
            self._unresolved += 1
        need = int(self._footprint(req.nmesh))
        t = C.c_int64()
        rc = self._lib.pp_serve_submit(self._h, C.c_void_p(rid), need, C.byref(t))
        if rc != 0:
            msg = self._lib.pp_last_error().decode()
            with self._lock:
                self._reqs.pop(rid)
                self._unresolved -= 1
            if rc == -6:   # PP_ERR_CAPACITY: could never be admitted
                now = time.perf_counter()
                fut.set_result(TileResult(req.tile_id, req.nmesh, arrival=req.arrival, start=now, done=now,
                                          error=f"AdmissionError: {msg}"))
                return fut
            self._L.check(rc)
        fut.ticket = t.value
        return fut

    def inject_fault(self, fut: Future):
        """Make a queued tile fail with PP_ERR_INJECTED instead of running
        (failure tests)."""
        self._L.check(self._lib.pp_serve_inject_fault(self._h, fut.ticket))

    def stats(self) -> dict:
        import ctypes as C
        s = (C.c_int64 * 7)()
        self._L.check(self._lib.pp_serve_stats(self._h, s))
        return dict(zip(("waiting", "in_flight", "max_in_flight", "finished", "failed", "reserved", "budget"),
                        list(s)))

    @property
    def max_in_flight(self) -> int:
        return self._max_in_flight if self._h is None else self.stats()["max_in_flight"]

    def close(self):
        with self._lock:
            if self._closed:
                return
            self._closed = True
        self._done_thread.join()
        self._max_in_flight = self.stats()["max_in_flight"]
        self._lib.pp_serve_destroy(self._h)
        self._h = None
        if self._runner is not None:
            for s in self._states:
                self._runner.release_slot(s)
            if self._own_runner:
                self._runner.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ callbacks
    def _call(self, user, ctx, slot, ticket, request):
        """A slot thread of the native server runs one tile."""
        rid = int(request or 0)
        req, _ = self._reqs[rid]
        try:
            res = self._run(self._states[slot], req)
        except Exception as e:  # the tile fails, the server keeps serving
            res = TileResult(req.tile_id, req.nmesh, error=f"{type(e).__name__}: {e}")
        self._reqs[rid] = (req, self._reqs[rid][1], res)
        return 0 if res.error is None else -1

    def _collect(self):
        """Resolve futures in completion order."""
        import ctypes as C
        t, st, rq = C.c_int64(), C.c_int(), C.c_void_p()
        times = (C.c_double * 4)()
        buf = C.create_string_buffer(1024)
        while True:
            with self._lock:
                if self._closed and self._unresolved == 0:
                    return
            rc = self._lib.pp_serve_next(self._h, 100, C.byref(t), C.byref(rq), C.byref(st), times)
            if rc == -8:   # PP_ERR_TIMEOUT
                continue
            self._L.check(rc)
            # the request id travels as the opaque request pointer
            entry = self._reqs.pop(int(rq.value or 0))
            req, fut = entry[0], entry[1]
            res = entry[2] if len(entry) > 2 else None
            if res is None:
                self._lib.pp_serve_message(self._h, t.value, buf, len(buf))
                name = self._L.ERRORS.get(st.value, str(st.value))
                res = TileResult(req.tile_id, req.nmesh, error=f"{name}: {buf.value.decode()}")
            res.arrival = req.arrival
            res.start, res.done, res.slot = self._t0 + times[1], self._t0 + times[2], int(times[3])
            with self._lock:
                self._unresolved -= 1
            fut.set_result(res)


def replay(server: TileServer, stream, time_scale: float = 1.0, on_done=None):
    """Submit (offset, request) pairs at their offsets; return the results in
    request order once all have finished.  on_done(result) is called as each
    tile finishes (from the completion thread)."""
    t0 = time.perf_counter()
    futs = []
    for off, req in stream:
        dt = t0 + off * time_scale - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        f = server.submit(req)
        if on_done is not None:
            f.add_done_callback(lambda fu: on_done(fu.result()))
        futs.append(f)
    return [f.result() for f in futs]


def summarize(results) -> dict:
    ok = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(0)
    wall = (max(r.done for r in ok) - min(r.arrival for r in ok)) if ok else 0.0
    stages = {}
    for r in ok:
        for k, v in r.stages.items():
            stages[k] = stages.get(k, 0.0) + v
    return {"tiles": len(results), "completed": len(ok), "failed": len(results) - len(ok),
            "mean_stage_s": {k: v / len(ok) for k, v in stages.items()} if ok else {},
            "tiles_per_s": len(ok) / wall if wall > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50)) * 1e3 if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) * 1e3 if len(lat) else None,
            "peaks": int(sum(r.npeaks for r in ok)), "halos": int(sum(r.nhalo for r in ok)),
            "wall_s": wall}


class GpuTileRunner:
    """The per-tile GPU pipeline on one slot: synthesis -> filter bank +
    stencil -> per-peak bookkeeping -> shell analysis -> record packing, all
    on the slot context's stream; the peak list and the per-peak batch stay
    on the device.

    split=True (c5 as run_multitile_split serves it): a request names a tile
    (it, jt, kt) of its nmesh's global field; the fields come from that
    field's coarse grid (created on first use, then resident on the device
    for the runner's life and shared by every slot) and the tile's residual
    (pp_split_tile_fields), nbuff = the split geometry's.  split=False: each
    request is a periodic tile of its own seed (pp_make_tile_fields)."""

    def __init__(self, device: int = 0, template: TileConfig = SERVE_TEMPLATE, table=None, tunables=None,
                 round_f32: bool = True, split: bool = False, ntile: int = SPLIT_NTILE,
                 coarse_factor: int = COARSE_FACTOR, split_seed: int = SPLIT_SEED, trim_n: int = 512):
        import threading as _th

        from . import collapse
        self.split = split
        self.ntile, self.coarse_factor, self.split_seed = ntile, coarse_factor, split_seed
        # after a tile of nmesh >= trim_n the slot returns its pooled memory
        # (the context's scratch and torch's cached blocks), so an idle slot
        # does not hold a large tile's working set against the other slots
        self.trim_n = trim_n
        self._coarse = {}
        self._coarse_lock = _th.Lock()
        self.device = device
        # run_multitile_split stores per-peak ZZon and fcrit as Float32
        # (MultiResolution.jl:963-967)
        self.round_f32 = round_f32
        self.template = template
        self.table = table if table is not None else collapse.sinusoidal_table()
        self.tunables = tunables

    def make_slot(self, handle=None):
        """A slot's state: its Context (a new one, or a non-owning wrapper of
        the native server's slot context `handle`) with the tables loaded,
        and a torch stream on the context's stream."""
        import torch
        from .api import Context
        ctx = Context(self.device) if handle is None else Context.from_handle(handle, self.device)
        ctx.upload_collapse_table(self.table.table, *self.table.bounds)
        ctx.build_shell_tables(self.template.nhunt)
        if self.tunables:
            ctx.tunables(**self.tunables)
        stream = torch.cuda.ExternalStream(ctx.stream_handle(), device=torch.device("cuda", self.device))
        return {"ctx": ctx, "stream": stream}

    @staticmethod
    def release_slot(state):
        if state is not None:
            state["ctx"].close()

    def split_geometry(self, nmesh: int):
        from . import split as S
        return S.geometry(nmesh, self.template.cap + 1, self.ntile, self.template.alatt,
                          coarse_factor=self.coarse_factor, seed=self.split_seed)

    def coarse(self, ctx, nmesh: int):
        """The resident coarse field of nmesh's split run (made on first use)."""
        from . import split as S
        with self._coarse_lock:
            c = self._coarse.get(nmesh)
            if c is None:
                c = S.CoarseField(ctx, self.split_geometry(nmesh))
                self._coarse[nmesh] = c
            return c

    def footprint(self, nmesh: int) -> int:
        from . import split as S
        return S.footprint_bytes(nmesh) if self.split else footprint_bytes(nmesh)

    def close(self):
        with self._coarse_lock:
            for c in self._coarse.values():
                c.close()
            self._coarse.clear()

    def run(self, state, req: TileRequest) -> TileResult:
        import torch
        from . import catalog, synth
        from ._lib import SHELL_PERSISTENT
        from .pipeline import TileDriver
        ctx, stream = state["ctx"], state["stream"]
        cfg = tile_config(req, self.template)
        if self.split:
            from . import split as S
            g = self.split_geometry(req.nmesh)
            tix = req.tile if req.tile is not None else (1, 1, 1)
            cfg = replace(cfg, nbuff=g.nbuff, tile_center=g.tile_center(*tix), seed=g.seed)
        drv = TileDriver(cfg)
        st = {}
        t = [time.perf_counter()]

        def lap(name):
            stream.synchronize()
            now = time.perf_counter()
            st[name] = now - t[0]
            t[0] = now

        with torch.cuda.stream(stream):
            if self.split:
                tile = S.make_split_tile(ctx, self.coarse(ctx, req.nmesh), *tix, ilpt=cfg.ilpt,
                                         want_lapd=cfg.ioutshear >= 1)
            else:
                tile = synth.make_tile(ctx, cfg.n, cfg.boxsize, cfg.seed, ilpt=cfg.ilpt,
                                       want_lapd=cfg.ioutshear >= 1)
            lap("synthesis")
            peaks = drv.find_peaks(ctx, tile.delta, device_out=True)
            lap("peak_find")
            bdev = drv.device_batch(ctx, peaks, round_f32=self.round_f32 and cfg.ievol == 1)
            lap("bookkeeping")
            out = TileResult(req.tile_id, req.nmesh, npeaks=len(bdev["i"]), stages=st)
            if out.npeaks == 0:
                out.records = np.zeros((0, 33), np.float32)
                return out
            mask = torch.zeros((cfg.n,) * 3, dtype=torch.int8, device=f"cuda:{self.device}")
            fields = (tile.delta, tile.eta, tile.eta2, tile.lapd)
            res = drv.analyse(ctx, fields, bdev, mask=mask, kernel=SHELL_PERSISTENT, device_out=True)
            lap("shell_analysis")
            rec, rmax = catalog.pack_halos(ctx, res, bdev, cfg.alatt, extended=True, z_max=cfg.z_max)
            out.records = rec.cpu().numpy()
            lap("records")
        out.RTHLmax = rmax
        out.nhalo = len(out.records)
        if self.trim_n and cfg.n >= self.trim_n:
            del tile, peaks, bdev, mask, fields, res, rec
            ctx.release_scratch()
            torch.cuda.empty_cache()
        return out
