"""Parity at the full light-cone tile: nmesh = 512, box 600 Mpc/h, 2LPT, the
25-filter bank 1-30 Mpc/h (about 2 M peaks), set up as the benchmark sets
it up.

Properties that hold at any size:
- the persistent form is deterministic: two runs give bit-identical
  catalogues and masks;
- the shell count m of every peak equals the host's geometric count
  (RadialShell.jl:289-295, 379-453), in both forms;
- the plain and persistent forms agree on every collapse decision and
  R_TH for >= 99.9% of peaks (FP32, different summation orders);
- every peak against the FP64 oracle: collapse decisions >= 99.99%,
  R_TH within one shell bin and M200m within 0.5% for >= 99.9% of halos,
  and the exclusion masks differ on < 0.1% of the masked cells (measured:
  100%, 100%, 100% and no cell).
"""
import json
import os
import time

import numpy as np
import pytest

from precision_study import compare as compare_catalogs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_c3_full_tile():
    import torch
    from oracle import oracle as O
    from peakpatch_hip import SHELL_PERSISTENT, SHELL_PLAIN, Context, collapse, synth
    from peakpatch_hip.pipeline import CONFIGS, TileDriver

    cfg = CONFIGS["c3"]
    assert cfg.n == 512
    ct = collapse.read_homeltab(os.path.join(ROOT, "tests", "golden", "HomelTab_fortran.dat"))
    ctx = Context(0)
    try:
        ctx.upload_collapse_table(ct.table, *ct.bounds)
        ctx.build_shell_tables(cfg.nhunt)
        tile = synth.make_tile(ctx, cfg.n, cfg.boxsize, cfg.seed, ilpt=cfg.ilpt, want_lapd=cfg.ioutshear >= 1)
        ctx.synchronize()
        drv = TileDriver(cfg)
        b = drv.batch(drv.find_peaks(ctx, tile.delta))
        npk = len(b)
        print(f"[c3] {npk} peaks", flush=True)
        assert npk > 1_000_000
        fields = (tile.delta, tile.eta, tile.eta2, tile.lapd)
        bdev = b.to_device()

        def run(form):
            mask = torch.zeros((cfg.n,) * 3, dtype=torch.int8, device="cuda")
            r = drv.analyse(ctx, fields, bdev, mask=mask, kernel=form, device_out=True)
            torch.cuda.synchronize()
            r = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in r.items()}
            return r, mask.cpu().numpy()

        p1, m1 = run(SHELL_PERSISTENT)
        p2, m2 = run(SHELL_PERSISTENT)
        assert p1["dump_counts"] == p2["dump_counts"]
        for f in p1:
            if f == "dump_counts":
                continue
            a, c = np.asarray(p1[f]), np.asarray(p2[f])
            assert a.shape == c.shape and a.tobytes() == c.tobytes(), f"persistent form not deterministic: {f}"
        assert np.array_equal(m1, m2)

        _, mshell = drv.gather_extent(b)
        assert np.array_equal(p1["nshells"], mshell.astype(np.int32))
        q, mq = run(SHELL_PLAIN)
        assert np.array_equal(q["nshells"], p1["nshells"])
        forms = compare_catalogs(q, p1, q["status"], p1["status"], rad_of(cfg.nhunt), cfg.alatt,
                                 cfg.cosmology.Om, cfg.cosmology.h)
        print("[c3] plain vs persistent", json.dumps({k: v for k, v in forms.items() if k != "fields"}), flush=True)
        assert forms["status_agreement"] >= 0.999
        assert forms["rthl_within_one_bin"] >= 0.999

        host = lambda t: t.detach().cpu().numpy()
        mask_c = np.zeros((cfg.n,) * 3, np.int8)
        grid = O.Grid(host(tile.delta), *(host(a) for a in tile.eta),
                      eta2=tuple(host(a) for a in tile.eta2) if tile.eta2 is not None else None,
                      mask=mask_c, lapd=host(tile.lapd) if tile.lapd is not None else None)
        n = cfg.n
        ipp = (b.i.astype(np.int64) + 1) + b.j.astype(np.int64) * n + b.k.astype(np.int64) * n * n
        t0 = time.perf_counter()
        ref, st = O.analyse_peaks(grid, ipp, b.ir2min, b.ZZon, b.Rf, b.fcrit, cfg.alatt, O.Table(ct.table, *ct.bounds),
                                  O.Shells(cfg.nhunt), cfg.nbuff_eff, cfg.rmax2rs, min(16, os.cpu_count() or 1))
        print(f"[c3] oracle over all peaks in {time.perf_counter() - t0:.1f} s", flush=True)
        c = compare_catalogs(p1, ref, p1["status"], st, rad_of(cfg.nhunt), cfg.alatt, cfg.cosmology.Om,
                             cfg.cosmology.h)
        both = (p1["status"] == 0) & (st == 0)
        c["rthl_exact"] = float((np.asarray(p1["RTHL"])[both] == ref["RTHL"][both]).mean())
        masked = (m1 != 0) | (mask_c != 0)
        c["mask_cells"] = int(masked.sum())
        c["mask_mismatch_frac"] = float(((m1 != 0) != (mask_c != 0)).sum() / max(1, masked.sum()))
        print("[c3] persistent vs oracle", json.dumps({k: v for k, v in c.items() if k != "fields"}), flush=True)
        assert c["halos_oracle"] > 10_000
        assert c["status_agreement"] >= 0.9999
        assert c["rthl_within_one_bin"] >= 0.999
        assert c["mass_within_0.5pct"] >= 0.999
        assert c["mask_mismatch_frac"] < 1e-3
        cq = compare_catalogs(q, ref, q["status"], st, rad_of(cfg.nhunt), cfg.alatt, cfg.cosmology.Om,
                              cfg.cosmology.h)
        cq["mask_mismatch_frac"] = float(((mq != 0) != (mask_c != 0)).sum() / max(1, masked.sum()))
        print("[c3] plain vs oracle", json.dumps({k: v for k, v in cq.items() if k != "fields"}), flush=True)
        assert cq["status_agreement"] >= 0.9999
        assert cq["rthl_within_one_bin"] >= 0.999
        assert cq["mass_within_0.5pct"] >= 0.999
        assert cq["mask_mismatch_frac"] < 1e-3
    finally:
        ctx.close()


def rad_of(nhunt):
    """Distinct shell radii sqrt(r^2) of the shell table out to nhunt."""
    from oracle import oracle as O
    return np.sqrt(np.unique(O.Shells(nhunt).r2).astype(np.float64))
