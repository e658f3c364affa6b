#!/usr/bin/env python
"""Benchmark of the halo-finding hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3]

Workload (default c3): one light-cone tile, nmesh=512, box 600 Mpc/h, 2LPT,
ioutshear=1, 25 log-spaced top-hat filters 1-30 Mpc/h, rmax2rs=2, observer at
the box corner, seed 42.  Set-up (untimed): the fields are synthesised on the
GPU, the filter bank + stencil finds the peaks, the host computes per-peak
redshift / fcrit.  A step is one shell analysis of every peak of the tile
(gather + virialisation walk + R_TH + mask), fields resident in HBM.

Prints one JSON line: shell-analysed peaks/s, the roofline of the dominant
kernel (the wave gather: 4 B x 5 fields x the cells it reads, against its
HIP-event time; the whole call's 4 B x F fields x cells figure is reported
beside it), the p50/p99 latency of a stream of tiles through the serving
loop, and the CPU baseline (the FP64 restatement of the reference algorithm,
OpenMP over peaks) timed on a bounded stratified sample of the same peaks.
"""
from __future__ import annotations

import argparse
import gc
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "peakpatch.jl_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "shell-analysed peaks/s at nmesh=512 + % HBM peak; p50/p99 tile latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def nfields(cfg) -> int:
    return 1 + 1 + 3 + (3 if cfg.ilpt >= 2 else 0) + (1 if cfg.ioutshear >= 1 else 0)


def load_table(ctx):
    from peakpatch_hip import collapse
    path = os.path.join(ROOT, "tests", "golden", "HomelTab_fortran.dat")
    ct = collapse.read_homeltab(path) if os.path.exists(path) else collapse.sinusoidal_table()
    ctx.upload_collapse_table(ct.table, *ct.bounds)
    return ct, os.path.basename(path) if os.path.exists(path) else "sinusoidal"


def time_form(drv, ctx, fields, bdev, mask, form, steps, warmup):
    """Device-resident steps: batch, fields, mask and results stay in HBM."""
    import torch
    # a collection of Python's cyclic garbage collector inside a timed step
    # (tens of ms with torch loaded) would be timed as the step's; collect
    # now, before the warm-up steps, and hold it off until the timed steps
    # are done
    gc.collect()
    gc.disable()
    walls, kern, stages = [], [], []
    res = None
    try:
        for _ in range(warmup):
            mask.zero_()
            torch.cuda.synchronize()
            drv.analyse(ctx, fields, bdev, mask=mask, kernel=form, device_out=True)
        for _ in range(steps):
            res = None  # the previous step's result arrays go back to the allocator before this step
            mask.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = drv.analyse(ctx, fields, bdev, mask=mask, kernel=form, device_out=True)
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t0)
            tm = ctx.last_timing()
            kern.append(tm["shell_ms"] / 1e3)
            stages.append(tm["stages_ms"])
    finally:
        gc.enable()
    res = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in res.items()}
    return res, walls, kern, stages


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, tile, b, ct, seconds: float, runs: int = 5):
    """Oracle (FP64 C restatement, OpenMP dynamic over peaks) on the tile's
    peaks: the median of `runs` timings on every CPU this process may use
    (a stratified sample sized to about `seconds` per run), and one run on a
    single core over a smaller stratified sample."""
    import numpy as np
    from oracle import oracle as O
    avail = len(os.sched_getaffinity(0))
    # OMP_NUM_THREADS states the CPU share a job is given where one is set
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(avail, share) if share > 0 else avail
    host = lambda t: t.detach().cpu().numpy()
    delta = host(tile.delta)
    eta = tuple(host(a) for a in tile.eta)
    eta2 = tuple(host(a) for a in tile.eta2) if tile.eta2 is not None else None
    lapd = host(tile.lapd) if tile.lapd is not None else None
    mask = np.zeros(delta.shape, np.int8)
    grid = O.Grid(delta, *eta, eta2=eta2, mask=mask, lapd=lapd)
    tab = O.Table(ct.table, *ct.bounds)
    sh = O.Shells(cfg.nhunt)
    n = cfg.n
    ipp = (b.i.astype(np.int64) + 1) + b.j.astype(np.int64) * n + b.k.astype(np.int64) * n * n

    def run(sel, nth):
        t0 = time.perf_counter()
        O.analyse_peaks(grid, ipp[sel], b.ir2min[sel], b.ZZon[sel], b.Rf[sel], b.fcrit[sel], cfg.alatt, tab, sh,
                        cfg.nbuff_eff, cfg.rmax2rs, nth)
        return time.perf_counter() - t0

    N = len(ipp)
    stride = max(1, N // 2000)
    t = run(np.arange(0, N, stride), threads)
    est_all = t * stride
    stride2 = max(1, int(math.ceil(est_all / seconds)))
    sel = np.arange(0, N, stride2)
    times = [run(sel, threads) for _ in range(runs)]
    t = float(np.median(times))
    # one core: a sample sized from the multi-core rate (about `seconds`)
    stride1 = max(1, int(math.ceil(N * t / len(sel) * threads / seconds)))
    sel1 = np.arange(0, N, stride1)
    t1 = run(sel1, 1)
    return {"value": len(sel) / t, "unit": "peaks/s", "cores": threads, "kind": "port",
            "sample": f"every {stride2}th peak of the same tile ({len(sel)} of {N}, all filters), median of "
                      f"{runs} runs ({min(times):.2f}-{max(times):.2f} s), FP64 oracle restating "
                      f"RadialShell.analyse_peak, OpenMP dynamic over peaks",
            "runs_s": [round(x, 3) for x in times], "host_cpus": avail, "cpu_model": cpu_model(),
            "one_core": {"value": len(sel1) / t1, "unit": "peaks/s",
                         "sample": f"every {stride1}th peak ({len(sel1)}), {t1:.1f} s"}}


def read_traffic(paths, head, kernel, tuned):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summaries (FETCH_SIZE + WRITE_SIZE, each scaled by the
    same counter on a copy kernel of known size), and the launch parameters
    the profiled command pinned (--tune): the summary profiled at the launch
    parameters this run chose if there is one, else the first; or (None, None)."""
    if head != "persistent":
        return None, None
    found = []
    for path in paths:
        if not os.path.exists(path):
            continue
        traffic, tune = None, None
        with open(path) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch[{kernel}]"):
                    traffic = float(line.split("=", 1)[1].split()[0])
                elif line.startswith("command (FETCH_SIZE)") and "--tune " in line:
                    kv = line.split("--tune ", 1)[1].split()[0]
                    tune = {k: int(v) for k, v in (x.split("=") for x in kv.split(","))}
        if traffic is not None:
            found.append((traffic, tune))
    if not found:
        return None, None
    return next((x for x in found if x[1] == tuned), found[0])


def serve_stream(args, ct, tuned):
    """The c5 serving workload: run_multitile_split light-cone tiles
    (coarse_factor 5) with nmesh drawn from {128, 256, 384, 512, 768} at the c2
    resolution, each generated from its nmesh's resident coarse field plus its
    own isolated residual, peak-found, bookkept, shell-analysed and packed on
    the GPU, on the native server's slot contexts with admission by device
    memory.  Two runs:
    args.serve_burst tiles submitted at once (the drain rate is the saturation
    capacity), then args.serve_tiles Poisson arrivals at args.serve_load
    times that capacity (or at args.serve_rate); latency = completion -
    arrival, queueing included."""
    from peakpatch_hip import serve
    runner = serve.GpuTileRunner(table=ct, tunables=tuned, split=True)
    n = args.serve_tiles
    done = [0]

    def progress(what):
        def cb(r):
            done[0] += 1
            if done[0] % 25 == 0 or r.error:
                log(f"[serve] {what}: {done[0]} tiles done (last nmesh {r.nmesh}, "
                    f"{'error: ' + r.error if r.error else f'{r.latency * 1e3:.0f} ms'})")
        return cb
    with serve.TileServer(slots=args.serve_slots, runner=runner, hbm_fraction=0.8) as srv:
        # warm the slots on every nmesh (the resident coarse fields, transform
        # plans, buffer pools), untimed
        warm = [serve.TileRequest(-1, m, tile=(1, 1, 1)) for m in serve.SERVE_NMESH for _ in range(args.serve_slots)]
        t0 = time.perf_counter()
        bad = [r.error for r in serve.replay(srv, [(0.0, r) for r in warm], on_done=progress("warm-up")) if r.error]
        if bad:
            raise RuntimeError(f"serving warm-up failed: {bad[0]}")
        log(f"[serve] warm-up ({len(warm)} tiles, coarse fields of {len(serve.SERVE_NMESH)} runs) "
            f"{time.perf_counter() - t0:.1f} s")
        done[0] = 0
        burst = serve.poisson_stream(args.serve_burst, 0.0, seed=5)
        cap = serve.summarize(serve.replay(srv, burst, on_done=progress("capacity run")))
        log(f"[serve] capacity {cap['tiles_per_s']:.2f} tiles/s over {cap['wall_s']:.1f} s")
        done[0] = 0
        rate = args.serve_rate if args.serve_rate > 0 else args.serve_load * cap["tiles_per_s"]
        stream = serve.poisson_stream(n, rate, seed=6)
        res = serve.replay(srv, stream, on_done=progress(f"Poisson {rate:.2f}/s"))
        st = srv.stats()
    geoms = {m: runner.split_geometry(m) for m in serve.SERVE_NMESH}
    runner.close()
    s = serve.summarize(res)
    mesh = [r.nmesh for _, r in stream]
    s.update({"workload": f"c5: {n} run_multitile_split tiles (coarse_factor {serve.COARSE_FACTOR}, "
                          f"{serve.SPLIT_NTILE}^3 tiles per nmesh, coarse fields resident), Poisson {rate:.3g}/s, "
                          f"nmesh from {serve.SERVE_NMESH}, alatt 250/256 Mpc/h, 20 filters 1-15 Mpc/h, 2LPT, "
                          f"light-cone",
              "split_geometry": {str(m): {"N": g.N, "M": g.M, "block": g.block, "nbuff": g.nbuff}
                                 for m, g in geoms.items()},
              "arrival_rate_per_s": rate,
              "load": rate / cap["tiles_per_s"] if cap["tiles_per_s"] > 0 else None,
              "capacity_tiles_per_s": cap["tiles_per_s"],
              "capacity_run": {k: cap[k] for k in ("tiles", "completed", "failed", "wall_s", "p50_ms", "p99_ms")},
              "slots": args.serve_slots, "max_in_flight": st["max_in_flight"],
              "admission": {"budget_bytes": st["budget"], "rule": "footprint within 0.8 of the free HBM at start, "
                            "and within the live free HBM while others run"},
              "nmesh_counts": {str(m): mesh.count(m) for m in serve.SERVE_NMESH}})
    if s["failed"] or cap["failed"]:
        s["errors"] = sorted({r.error for r in res if r.error})
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--form", default="auto", choices=["auto", "plain", "persistent"])
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain per-peak form")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch parameters")
    ap.add_argument("--tune", default="", help="pin launch parameters, e.g. small_m=12,chunk_shells=128 "
                    "(implies --no-autotune)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="before the timed steps, run one elementwise kernel over a GIB-sized float32 buffer "
                         "(known traffic, for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--serve-tiles", type=int, default=200,
                    help="tiles in the serving latency run (nmesh in {128..768}); 0 skips serving")
    ap.add_argument("--serve-burst", type=int, default=60,
                    help="tiles submitted at once to measure the serving capacity (drain rate)")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="Poisson arrival rate of the latency run, as a fraction of the measured capacity")
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="Poisson arrival rate in tiles/s (overrides --serve-load when > 0)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-only", action="store_true", help="run only the serving stream (a test of it)")
    ap.add_argument("--traffic-file", nargs="+",
                    default=[os.path.join(ROOT, "profiles", f) for f in ("k_gather_c3_pmc.txt",
                                                                         "k_gather_c3_pmc_gwc24.txt")],
                    help="rocprofv3 --pmc summaries whose 'traffic_bytes_per_launch' fills roofline.traffic "
                         "(the one profiled at this run's launch parameters, else the first)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(tiles would shard one per GPU; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from peakpatch_hip import SHELL_PERSISTENT, SHELL_PLAIN, Context, PPError, synth
    from peakpatch_hip.pipeline import CONFIGS, TileDriver

    cfg = CONFIGS[args.config]
    ctx = Context(0)
    ct, table_name = load_table(ctx)
    if args.serve_only:
        print(json.dumps({"serving": serve_stream(args, ct, ctx.tunables())}), flush=True)
        return
    ctx.build_shell_tables(cfg.nhunt)
    t0 = time.perf_counter()
    tile = synth.make_tile(ctx, cfg.n, cfg.boxsize, cfg.seed, ilpt=cfg.ilpt, want_lapd=cfg.ioutshear >= 1)
    ctx.synchronize()
    torch.cuda.synchronize()
    log(f"[bench] {cfg.describe()}; fields synthesised in {time.perf_counter() - t0:.1f} s")
    drv = TileDriver(cfg)
    t0 = time.perf_counter()
    peaks = drv.find_peaks(ctx, tile.delta)
    pf_wall = time.perf_counter() - t0
    pf_ms = ctx.last_timing()["peakfind_ms"]
    b = drv.batch(peaks)
    log(f"[bench] filter bank + stencil: {len(peaks)} peaks, {pf_ms:.1f} ms on device ({pf_wall:.2f} s wall)")
    fields = (tile.delta, tile.eta, tile.eta2, tile.lapd)
    mask = torch.zeros((cfg.n,) * 3, dtype=torch.int8, device="cuda")
    bdev = b.to_device()
    prop = torch.cuda.get_device_properties(0)
    log(f"[bench] device {prop.name}: {prop.multi_processor_count} CUs, {prop.total_memory / 2**30:.0f} GiB")

    if args.tune:
        ctx.tunables(**{k: int(v) for k, v in (kv.split("=") for kv in args.tune.split(","))})
        args.no_autotune = True
    tuned = ctx.tunables()
    if not args.no_autotune and args.form in ("auto", "persistent"):
        t0 = time.perf_counter()
        c = cfg
        tuned = ctx.autotune_shell(fields[0], fields[1], bdev["i"], bdev["j"], bdev["k"], bdev["Rf"], bdev["ir2min"],
                                   bdev["ZZon"], bdev["fcrit"], c.alatt, eta2=fields[2] if c.ilpt >= 2 else None,
                                   lapd=fields[3] if c.ioutshear >= 1 else None, rmax2rs=c.rmax2rs)
        log(f"[bench] autotuned launch parameters in {time.perf_counter() - t0:.1f} s: {tuned}")

    if args.copy_calibration > 0:
        nel = int(args.copy_calibration * 2**30) // 4
        src = torch.ones(nel, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.add(src, 1.0, out=dst)  # an elementwise kernel (a same-type copy_ may go to the DMA engine)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: elementwise add, {nel * 4} bytes read, {nel * 4} bytes written")
        del src, dst

    forms = []
    if args.form in ("auto", "persistent"):
        forms.append(("persistent", SHELL_PERSISTENT))
    if args.form == "plain" or (args.form == "auto" and not args.no_plain):
        forms.append(("plain", SHELL_PLAIN))
    timings = {}
    results = {}
    for name, form in forms:
        try:
            res, walls, kern, stages = time_form(drv, ctx, fields, bdev, mask, form, args.steps, args.warmup)
        except PPError as e:
            log(f"[bench] {name} form unavailable: {e}")
            continue
        timings[name] = (walls, kern, stages)
        results[name] = res
        log(f"[bench] {name}: wall {np.median(walls) * 1e3:.1f} ms/step, device {np.median(kern) * 1e3:.1f} ms, "
            f"stages {stages[-1]}, collapsed {(res['RTHL'] > 0).sum()}, dumps {res['dump_counts']}")
    head = "persistent" if "persistent" in timings else "plain"
    walls, kern, stages = timings[head]
    npk = len(b)
    ms_step = float(np.mean(walls)) * 1e3
    k_s = float(np.mean(kern))
    cells, mshell = drv.gather_extent(b)
    F = nfields(cfg)
    stage_ms = {k: float(np.mean([st[k] for st in stages])) for k in stages[0]}
    if head == "persistent":
        # dominant kernel: the wave gather (k_gather) over the peaks above the
        # small-peak threshold; it reads 5 fields per gathered cell (delta,
        # its transpose, eta x/y/z; eta2 and lapd are read by the finish pass
        # inside R_TH only), so its bytes are 4 B x 5 x the cells it reads
        big = mshell > tuned["small_m"]
        dom_bytes = float(4 * 5 * cells[big].sum())
        dom_s = stage_ms["gather"] / 1e3
        dom_name = "k_gather"
    else:
        dom_bytes = float(4 * F * cells.sum())
        dom_s = k_s
        dom_name = "k_shell_plain"
    achieved = dom_bytes / dom_s / 1e9
    out = {
        "metric": METRIC,
        "value": npk / (ms_step / 1e3),
        "unit": "peaks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "step_ms": [round(w * 1e3, 3) for w in walls],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg.describe(), "npeaks": npk, "kernel_form": head,
                   "collapse_table": table_name, "fields_F": F, "cells_gathered": int(cells.sum())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom_name,
                     "kernel_ms": dom_s * 1e3, "algorithmic_bytes": dom_bytes},
        "kernel_ms_per_step": k_s * 1e3,
        "stages_ms": stage_ms,
        "tunables": tuned,
        "shell_algorithmic_GBps": float(4 * F * cells.sum()) / k_s / 1e9,
        "kernel_peaks_per_s": npk / k_s,
        "peakfind_ms": pf_ms,
        "halos": int((results[head]["RTHL"] > 0).sum()),
    }
    if "plain" in timings and head != "plain":
        pw, pk, _ = timings["plain"]
        out["plain_kernel_ms"] = float(np.mean(pk)) * 1e3
        out["plain_peaks_per_s"] = npk / float(np.mean(pw))
        out["speedup_vs_plain_kernel"] = float(np.mean(pk)) / k_s
    traffic, traffic_tune = read_traffic(args.traffic_file, head, dom_name, tuned)
    out["roofline"]["traffic"] = traffic
    if traffic is not None:
        # the profiled run's launch parameters; the traffic is that run's
        out["roofline"]["traffic_launch_parameters"] = traffic_tune
        out["roofline"]["traffic_same_launch_parameters"] = traffic_tune == tuned
    if args.serve_tiles > 0:
        out["serving"] = serve_stream(args, ct, tuned)
        log(f"[bench] serving: {out['serving']}")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, tile, b, ct, args.cpu_seconds)
        log(f"[bench] cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
