"""Precision of the GPU catalogue against the FP64 oracle on the same tile:
per-catalogue-field errors, collapse agreement, R_TH within one shell bin
and M200m within 0.5%, for the FP32-stored fields the kernels read and for
an FP16-stored variant (fields rounded to half precision and back).

    python tests/precision_study.py --config c4 [--n 768] [--no-fp16]

Test infrastructure (it runs the oracle); tests/test_gpu_precision.py runs
it at a reduced box with the c4 resolution.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "peakpatch.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

FIELDS = ("RTHL", "Srb", "Sbar", "Sbar2", "strain", "gradpk", "gradpkrf", "Fbarx", "e_v", "p_v",
          "zvir_half", "gradpkf", "d2F")


def shell_bin(rad: np.ndarray, r: np.ndarray) -> np.ndarray:
    """Width of the shell bin holding radius r: rad(s+1) - rad(s) with
    rad(s) <= r < rad(s+1)."""
    s = np.clip(np.searchsorted(rad, r, side="right") - 1, 0, len(rad) - 2)
    return rad[s + 1] - rad[s]


def compare(gpu: dict, ref: dict, st_gpu, st_ref, rad, alatt: float, Om: float, h: float) -> dict:
    from peakpatch_hip.catalog import halo_mass
    both = (st_gpu == 0) & (st_ref == 0)
    out = {"peaks": int(len(st_ref)), "halos_oracle": int((st_ref == 0).sum()), "halos_gpu": int((st_gpu == 0).sum()),
           "status_agreement": float((st_gpu == st_ref).mean()) if len(st_ref) else 1.0}
    if not both.any():
        return out
    rg, rr = np.asarray(gpu["RTHL"])[both], np.asarray(ref["RTHL"])[both]
    out["rthl_within_one_bin"] = float((np.abs(rg - rr) <= shell_bin(rad, rr)).mean())
    mg, mr = halo_mass(rg * alatt, Om, h), halo_mass(rr * alatt, Om, h)
    mrel = np.abs(mg / mr - 1.0)
    out["mass_within_0.5pct"] = float((mrel <= 5e-3).mean())
    out["mass_rel_err_p50"] = float(np.median(mrel))
    out["mass_rel_err_p99"] = float(np.percentile(mrel, 99))
    per = {}
    for f in FIELDS:
        g = np.asarray(gpu[f])[both].reshape(both.sum(), -1).astype(np.float64)
        o = np.asarray(ref[f])[both].reshape(both.sum(), -1).astype(np.float64)
        ok = np.isfinite(g) & np.isfinite(o)
        scale = np.nanmax(np.abs(np.where(ok, o, np.nan)), axis=0, keepdims=True) + 1e-300
        err = np.where(ok, np.abs(g - o) / scale, 0.0)
        rel = np.where(ok & (np.abs(o) > 1e-3 * scale), np.abs(g - o) / np.maximum(np.abs(o), 1e-300), 0.0)
        per[f] = {"max_err_over_scale": float(err.max()), "p99_err_over_scale": float(np.percentile(err, 99)),
                  "p50_rel": float(np.median(rel)), "p99_rel": float(np.percentile(rel, 99)),
                  "nan_pattern_equal": bool(np.array_equal(np.isnan(g), np.isnan(o)))}
    out["fields"] = per
    return out


def study(cfg_name: str = "c4", n: int | None = None, fp16: bool = True, threads: int = 16, seed: int | None = None,
          table_path: str | None = None) -> dict:
    import torch
    from oracle import oracle as O
    from peakpatch_hip import Context, collapse, synth
    from peakpatch_hip._lib import SHELL_PERSISTENT
    from peakpatch_hip.pipeline import CONFIGS, TileDriver

    cfg = CONFIGS[cfg_name]
    if n and n != cfg.n:
        # same resolution (alatt) in a smaller box
        from dataclasses import replace
        cfg = replace(cfg, n=n, boxsize=cfg.alatt * n, name=f"{cfg.name}@{n}")
    if seed is not None:
        from dataclasses import replace
        cfg = replace(cfg, seed=seed)
    table_path = table_path or os.path.join(ROOT, "tests", "golden", "HomelTab_fortran.dat")
    ct = collapse.read_homeltab(table_path)
    ctx = Context(0)
    ctx.upload_collapse_table(ct.table, *ct.bounds)
    ctx.build_shell_tables(cfg.nhunt)
    t0 = time.perf_counter()
    tile = synth.make_tile(ctx, cfg.n, cfg.boxsize, cfg.seed, ilpt=cfg.ilpt, want_lapd=cfg.ioutshear >= 1)
    drv = TileDriver(cfg)
    peaks = drv.find_peaks(ctx, tile.delta)
    b = drv.batch(peaks)
    fields = (tile.delta, tile.eta, tile.eta2, tile.lapd)
    host = lambda t: t.detach().cpu().numpy()
    hf = (host(tile.delta), tuple(host(a) for a in tile.eta), tuple(host(a) for a in tile.eta2) if tile.eta2 else None,
          host(tile.lapd) if tile.lapd is not None else None)
    t_setup = time.perf_counter() - t0

    def gpu_run(flds):
        mask = torch.zeros((cfg.n,) * 3, dtype=torch.int8, device="cuda")
        r = drv.analyse(ctx, flds, b.to_device(), mask=mask, kernel=SHELL_PERSISTENT, device_out=True)
        return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in r.items()}

    g32 = gpu_run(fields)
    grid = O.Grid(hf[0], *hf[1], eta2=hf[2], mask=np.zeros(hf[0].shape, np.int8), lapd=hf[3])
    ipp = (b.i.astype(np.int64) + 1) + b.j.astype(np.int64) * cfg.n + b.k.astype(np.int64) * cfg.n * cfg.n
    t1 = time.perf_counter()
    ref, st = O.analyse_peaks(grid, ipp, b.ir2min, b.ZZon, b.Rf, b.fcrit, cfg.alatt, O.Table(ct.table, *ct.bounds),
                              O.Shells(cfg.nhunt), cfg.nbuff_eff, cfg.rmax2rs, threads)
    t_oracle = time.perf_counter() - t1
    sh = O.Shells(cfg.nhunt)
    rad = np.sqrt(np.unique(sh.r2).astype(np.float64))
    cosmo = cfg.cosmology
    out = {"workload": cfg.describe(), "setup_s": t_setup, "oracle_s": t_oracle, "oracle_threads": threads,
           "fp32_fields": compare(g32, ref, g32["status"], st, rad, cfg.alatt, cosmo.Om, cosmo.h)}
    if fp16:
        half = lambda t: None if t is None else t.to(torch.float16).to(torch.float32)
        f16 = (half(tile.delta), tuple(half(a) for a in tile.eta),
               tuple(half(a) for a in tile.eta2) if tile.eta2 else None, half(tile.lapd))
        g16 = gpu_run(f16)
        c16 = compare(g16, ref, g16["status"], st, rad, cfg.alatt, cosmo.Om, cosmo.h)
        c16["accepted"] = bool(c16.get("rthl_within_one_bin", 0.0) >= 0.995)
        out["fp16_stored_fields"] = c16
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c4")
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--threads", type=int, default=16)
    args = ap.parse_args()
    r = study(args.config, args.n or None, fp16=not args.no_fp16, threads=args.threads)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
