"""The workload configurations end to end on the GPU against the oracle.

Peak lists (c1, c2, c3), in two steps:
1. the stencil is bit-exact: the oracle's find_peaks, run filter by filter
   with the shared mask on the GPU's own smoothed fields (pp_smooth_field on
   the GPU's rfft of the same delta), reproduces the GPU filter bank's peak
   list -- positions, filter, order, delta_s, lapd_s -- and final mask bit
   for bit;
2. the filter bank against the FP64 oracle (FP64 irfft of the oracle's
   delta_k, same window code): filter by filter, from the same incoming mask,
   every peak found in one field and not the other is a near-tie, i.e. its
   deciding comparison (against fcrit or a neighbour) is within twice the
   largest difference between the two fields.  The count of such peaks is
   printed per configuration.

Catalogues (c1, c2): every peak through both GPU forms against the FP64
oracle (collapse decisions, R_TH within one shell bin, M200m within 0.5%,
exclusion masks).

c4 (nmesh 768): properties that hold at any size (determinism, shell counts
equal to the geometric count, plain and persistent forms agreeing), plus the
oracle on a stratified sample of the peaks.
"""
import json
import math
import os
import time

import numpy as np
import pytest

from precision_study import compare as compare_catalogs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
THREADS = min(16, len(os.sched_getaffinity(0)))


def _table():
    from peakpatch_hip import collapse
    return collapse.read_homeltab(os.path.join(ROOT, "tests", "golden", "HomelTab_fortran.dat"))


def _tile(ctx, cfg):
    import torch
    from peakpatch_hip import synth
    t = synth.make_tile(ctx, cfg.n, cfg.boxsize, cfg.seed, ilpt=cfg.ilpt, want_lapd=cfg.ioutshear >= 1)
    torch.cuda.synchronize()
    return t


def _host(t):
    return t.detach().cpu().numpy()


def _irfft64(dk, n, box, Rf, wsmooth):
    """The oracle's smooth_field with a threaded FP64 inverse transform."""
    import scipy.fft as sf
    from oracle import oracle as O
    wk = O.window_multiply(dk, n, box, Rf, wsmooth).astype(np.complex128)
    return sf.irfftn(wk, s=(n, n, n), axes=(0, 1, 2), workers=THREADS)


def _margins(s, cells, fcrit):
    """min(ff - fcrit, min over the 26 neighbours of ff - neighbour) at cells
    (k, j, i), in float64: >= 0 exactly where the stencil accepts the cell
    (masks aside)."""
    k, j, i = cells
    ff = s[k, j, i].astype(np.float64)
    m = ff - np.float64(np.float32(fcrit))
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dz or dy or dx:
                    m = np.minimum(m, ff - s[k + dz, j + dy, i + dx].astype(np.float64))
    return m


def peak_list_parity(ctx, cfg, delta_h, delta_dev):
    import scipy.fft as sf
    from oracle import oracle as O
    from peakpatch_hip.pipeline import TileDriver
    n, box, nb = cfg.n, cfg.boxsize, cfg.nbuff_eff
    drv = TileDriver(cfg)
    fcrit = drv.filter_fcrit()
    mask_gpu = np.zeros((n, n, n), np.int8)
    pk = drv.find_peaks(ctx, delta_dev, mask=mask_gpu)
    dk_g = ctx.rfft(delta_h)
    dk_o = sf.rfftn(delta_h.astype(np.float64), axes=(0, 1, 2), workers=THREADS).astype(np.complex64)
    mask_a = np.zeros((n, n, n), np.int8)
    lists = []
    report = {"config": cfg.name, "peaks": len(pk), "filters": []}
    worst = 0.0
    for f, R in enumerate(cfg.filters):
        s_g = ctx.smooth_field(dk_g, n, box, float(R), cfg.wsmooth)
        before = mask_a.copy()
        i, j, k, v = O.find_peaks(s_g, mask_a, nb, float(fcrit[f]))
        lap = None
        if cfg.ioutshear >= 1:
            l_g = ctx.smooth_field(dk_g, n, box, float(R), 3)
            lap = l_g[k - 1, j - 1, i - 1]
        lists.append((np.full(len(i), f, np.int32), i - 1, j - 1, k - 1, v, lap))
        # the FP64 bank from the same incoming mask
        s_o = _irfft64(dk_o, n, box, float(R), cfg.wsmooth).astype(np.float32)
        mo = before.copy()
        io, jo, ko, _ = O.find_peaks(s_o, mo, nb, float(fcrit[f]))
        a = set(zip(i.tolist(), j.tolist(), k.tolist()))
        b = set(zip(io.tolist(), jo.tolist(), ko.tolist()))
        diff = sorted(a ^ b)
        eps = float(np.abs(s_g - s_o).max())
        if diff:
            d = np.array(diff, np.int64) - 1
            mg = _margins(s_o, (d[:, 2], d[:, 1], d[:, 0]), fcrit[f])
            ratio = float(np.abs(mg).max() / (2 * eps))
            worst = max(worst, ratio)
            assert ratio <= 1.0 + 1e-6, (cfg.name, f, diff[:5], mg[:5], eps)
        report["filters"].append({"Rf": round(float(R), 4), "gpu": len(a), "fp64": len(b), "differ": len(diff),
                                  "max_field_diff": eps})
    cat = [np.concatenate([l[q] for l in lists]) for q in range(5)]
    assert np.array_equal(pk.filt, cat[0]) and np.array_equal(pk.i, cat[1])
    assert np.array_equal(pk.j, cat[2]) and np.array_equal(pk.k, cat[3])
    assert np.array_equal(pk.delta_s, cat[4])
    if cfg.ioutshear >= 1:
        assert np.array_equal(pk.lapd_s, np.concatenate([l[5] for l in lists]))
    assert np.array_equal(mask_gpu, mask_a)
    report["differ_total"] = int(sum(x["differ"] for x in report["filters"]))
    report["fp64_total"] = int(sum(x["fp64"] for x in report["filters"]))
    report["worst_margin_over_2eps"] = worst
    print(f"[{cfg.name}] peak lists", json.dumps(report), flush=True)
    return pk, report


@pytest.mark.parametrize("name", ["c1", "c2", "c3"])
def test_peak_lists(name):
    from peakpatch_hip import Context
    from peakpatch_hip.pipeline import CONFIGS
    cfg = CONFIGS[name]
    with Context(0) as ctx:
        t = _tile(ctx, cfg)
        peak_list_parity(ctx, cfg, _host(t.delta), t.delta)


def _oracle_catalogue(cfg, tile, b, ct):
    from oracle import oracle as O
    mask = np.zeros((cfg.n,) * 3, np.int8)
    grid = O.Grid(_host(tile.delta), *(_host(a) for a in tile.eta),
                  eta2=tuple(_host(a) for a in tile.eta2) if tile.eta2 is not None else None,
                  mask=mask, lapd=_host(tile.lapd) if tile.lapd is not None else None)
    n = cfg.n
    ipp = (b.i.astype(np.int64) + 1) + b.j.astype(np.int64) * n + b.k.astype(np.int64) * n * n
    ref, st = O.analyse_peaks(grid, ipp, b.ir2min, b.ZZon, b.Rf, b.fcrit, cfg.alatt, O.Table(ct.table, *ct.bounds),
                              O.Shells(cfg.nhunt), cfg.nbuff_eff, cfg.rmax2rs, THREADS)
    return ref, st, mask


def rad_of(nhunt):
    from oracle import oracle as O
    return np.sqrt(np.unique(O.Shells(nhunt).r2).astype(np.float64))


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_catalogue_vs_oracle(name):
    import torch
    from peakpatch_hip import SHELL_PERSISTENT, SHELL_PLAIN, Context
    from peakpatch_hip.pipeline import CONFIGS, TileDriver
    cfg = CONFIGS[name]
    ct = _table()
    with Context(0) as ctx:
        ctx.upload_collapse_table(ct.table, *ct.bounds)
        ctx.build_shell_tables(cfg.nhunt)
        tile = _tile(ctx, cfg)
        drv = TileDriver(cfg)
        b = drv.batch(drv.find_peaks(ctx, tile.delta))
        fields = (tile.delta, tile.eta, tile.eta2, tile.lapd)
        ref, st, mask_c = _oracle_catalogue(cfg, tile, b, ct)
        _, mshell = drv.gather_extent(b)
        for form, fname in ((SHELL_PERSISTENT, "persistent"), (SHELL_PLAIN, "plain")):
            mask = torch.zeros((cfg.n,) * 3, dtype=torch.int8, device="cuda")
            r = drv.analyse(ctx, fields, b, mask=mask, kernel=form)
            m = mask.cpu().numpy()
            assert np.array_equal(r["nshells"], mshell.astype(np.int32))
            c = compare_catalogs(r, ref, r["status"], st, rad_of(cfg.nhunt), cfg.alatt, cfg.cosmology.Om,
                                 cfg.cosmology.h)
            both = (r["status"] == 0) & (st == 0)
            c["rthl_exact"] = float((np.asarray(r["RTHL"])[both] == ref["RTHL"][both]).mean())
            c["mask_cells"] = int(((m != 0) | (mask_c != 0)).sum())
            c["mask_differ"] = int(((m != 0) != (mask_c != 0)).sum())
            print(f"[{name}] {fname} vs oracle", json.dumps({k: v for k, v in c.items() if k != "fields"}),
                  flush=True)
            assert c["halos_oracle"] > 100
            assert c["status_agreement"] >= 0.999
            assert c["rthl_within_one_bin"] >= 0.995
            assert c["mass_within_0.5pct"] >= 0.995
            assert c["mask_differ"] <= 1e-3 * c["mask_cells"]


def test_c4_nmesh768():
    """c4 at its stated size: 250 Mpc/h, nmesh 768, 20 filters 1-15 Mpc/h,
    2LPT, z = 0.5."""
    import torch
    from peakpatch_hip import SHELL_PERSISTENT, SHELL_PLAIN, Context
    from peakpatch_hip.pipeline import CONFIGS, TileDriver
    cfg = CONFIGS["c4"]
    assert cfg.n == 768
    ct = _table()
    with Context(0) as ctx:
        ctx.upload_collapse_table(ct.table, *ct.bounds)
        ctx.build_shell_tables(cfg.nhunt)
        tile = _tile(ctx, cfg)
        drv = TileDriver(cfg)
        b = drv.batch(drv.find_peaks(ctx, tile.delta))
        print(f"[c4] {len(b)} peaks, nhunt {cfg.nhunt}", flush=True)
        assert len(b) > 100_000
        fields = (tile.delta, tile.eta, tile.eta2, tile.lapd)
        bdev = b.to_device()

        def run(form):
            mask = torch.zeros((cfg.n,) * 3, dtype=torch.int8, device="cuda")
            r = drv.analyse(ctx, fields, bdev, mask=mask, kernel=form, device_out=True)
            torch.cuda.synchronize()
            return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in r.items()}, mask

        p1, m1 = run(SHELL_PERSISTENT)
        p2, m2 = run(SHELL_PERSISTENT)
        for f in p1:
            if f != "dump_counts":
                assert np.asarray(p1[f]).tobytes() == np.asarray(p2[f]).tobytes(), f
        assert torch.equal(m1, m2)
        del m2
        _, mshell = drv.gather_extent(b)
        assert np.array_equal(p1["nshells"], mshell.astype(np.int32))
        q, _ = run(SHELL_PLAIN)
        assert np.array_equal(q["nshells"], p1["nshells"])
        rad = rad_of(cfg.nhunt)
        forms = compare_catalogs(q, p1, q["status"], p1["status"], rad, cfg.alatt, cfg.cosmology.Om,
                                 cfg.cosmology.h)
        print("[c4] plain vs persistent", json.dumps({k: v for k, v in forms.items() if k != "fields"}), flush=True)
        assert forms["status_agreement"] >= 0.999 and forms["rthl_within_one_bin"] >= 0.999
        # the oracle on a stratified sample (every k-th peak in the
        # filter-ordered list covers every filter radius)
        from oracle import oracle as O
        stride = max(1, len(b) // 100000)
        sel = np.arange(0, len(b), stride)
        grid = O.Grid(_host(tile.delta), *(_host(a) for a in tile.eta), eta2=tuple(_host(a) for a in tile.eta2),
                      lapd=_host(tile.lapd))
        n = cfg.n
        ipp = (b.i[sel].astype(np.int64) + 1) + b.j[sel].astype(np.int64) * n + b.k[sel].astype(np.int64) * n * n
        t0 = time.perf_counter()
        ref, st = O.analyse_peaks(grid, ipp, b.ir2min[sel], b.ZZon[sel], b.Rf[sel], b.fcrit[sel], cfg.alatt,
                                  O.Table(ct.table, *ct.bounds), O.Shells(cfg.nhunt), cfg.nbuff_eff, cfg.rmax2rs,
                                  THREADS)
        sub = {k: np.asarray(v)[sel] for k, v in p1.items() if k != "dump_counts"}
        c = compare_catalogs(sub, ref, sub["status"], st, rad, cfg.alatt, cfg.cosmology.Om, cfg.cosmology.h)
        print(f"[c4] persistent vs oracle on {len(sel)} peaks ({time.perf_counter() - t0:.1f} s)",
              json.dumps({k: v for k, v in c.items() if k != "fields"}), flush=True)
        assert c["halos_oracle"] > 100
        assert c["status_agreement"] >= 0.999
        assert c["rthl_within_one_bin"] >= 0.995
        assert c["mass_within_0.5pct"] >= 0.995
