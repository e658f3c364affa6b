"""Per-peak bookkeeping on the device (pp_bookkeep_peaks) against the host
loop of TileDriver.batch (src/Pipeline.jl:217-231, 257-271;
src/MultiResolution.jl:959-988).

Both read the same peak list: find_peaks(device_out=True) must first equal
the host-output peak list bit for bit.  Then, per kept row: i, j, k, Rf,
ir2min, FcRf, d2Rf and the keep decision bit-exact; x, y, z bit-exact (same
Float64 operations); ZZon bit-exact at fixed z and within one ulp in
light-cone mode (1 + z, z from the same chi -> z bisection and
interpolation at an observer distance whose square root may differ by an
ulp); D, fcrit and v2fac within 1e-13 relative (the
device's 4-point Lagrange weights group their divisions differently from the
host's, and pow may differ by an ulp).  With round_f32 the Float32 rounding
of ZZon and fcrit must match the host's rounding except where the Float64
values sit within 1e-13 of a Float32 rounding boundary.
"""
from dataclasses import replace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from peakpatch_hip.api import Context
    return Context(0)


def _case(name):
    from peakpatch_hip import pipeline as P
    base = P.CONFIGS
    if name == "fixed_z":
        return P.scaled(base["c2"], 128)
    if name == "lightcone":
        return P.scaled(base["c3"], 128)
    if name == "lightcone_cut":
        # z_max inside the tile's redshift range, so some peaks are dropped
        return replace(P.scaled(base["c3"], 128), z_max=0.2)
    raise KeyError(name)


def _peaks(ctx, cfg):
    import torch
    from peakpatch_hip import synth
    from peakpatch_hip.pipeline import TileDriver
    t = synth.make_tile(ctx, cfg.n, cfg.boxsize, cfg.seed, ilpt=cfg.ilpt, want_lapd=cfg.ioutshear >= 1)
    drv = TileDriver(cfg)
    host = drv.find_peaks(ctx, t.delta)
    dev = drv.find_peaks(ctx, t.delta, device_out=True)
    torch.cuda.synchronize()
    return drv, host, dev


@pytest.mark.parametrize("name", ["fixed_z", "lightcone", "lightcone_cut"])
@pytest.mark.parametrize("round_f32", [False, True])
def test_bookkeep_matches_host(ctx, name, round_f32):
    cfg = _case(name)
    drv, host, dev = _peaks(ctx, cfg)
    print(f"\n[bookkeep] {name} round_f32={round_f32} n={cfg.n} peaks={len(host)}", flush=True)
    assert len(host) > 100
    for f in ("i", "j", "k", "filt", "delta_s"):
        np.testing.assert_array_equal(getattr(dev, f).cpu().numpy(), getattr(host, f), err_msg=f)
    if host.lapd_s is not None:
        np.testing.assert_array_equal(dev.lapd_s.cpu().numpy(), host.lapd_s)

    hb = drv.batch(host, round_f32=round_f32)
    db = {f: t.cpu().numpy() for f, t in drv.device_batch(ctx, dev, round_f32=round_f32).items()}
    if name == "lightcone_cut":
        assert 0 < len(hb) < len(host), "z_max must drop some peaks and keep some"
    assert len(db["i"]) == len(hb)
    for f in ("i", "j", "k", "Rf", "ir2min", "x", "y", "z", "FcRf", "d2Rf"):
        np.testing.assert_array_equal(db[f], getattr(hb, f), err_msg=f)
    for f in ("D", "v2fac"):
        np.testing.assert_allclose(db[f], getattr(hb, f), rtol=1e-13, atol=0, err_msg=f)
    if cfg.ilpt >= 2:
        assert np.all(db["v2fac"] > 0)
    if not round_f32:
        if cfg.ievol == 1:
            # 1 + z with z interpolated at the observer distance: the device's
            # sqrt and the host's agree to one ulp
            np.testing.assert_allclose(db["ZZon"], hb.ZZon, rtol=3e-16, atol=0)
        else:
            np.testing.assert_array_equal(db["ZZon"], hb.ZZon)
        np.testing.assert_allclose(db["fcrit"], hb.fcrit, rtol=1e-13, atol=0)
    else:
        for f in ("ZZon", "fcrit"):
            a, b = db[f], getattr(hb, f)
            assert np.all(a.astype(np.float32).astype(np.float64) == a), f"{f} not Float32-valued"
            off = np.flatnonzero(a != b)
            # a difference is a Float32 rounding of values within 1e-13 of a boundary: one ulp apart
            if len(off):
                ulp = np.spacing(np.abs(b[off]).astype(np.float32)).astype(np.float64)
                assert np.all(np.abs(a[off] - b[off]) <= ulp), f
            assert len(off) <= max(1, len(a) // 100_000), f"{f}: {len(off)} rows differ"


def test_bookkeep_empty(ctx):
    from peakpatch_hip.api import PeakList
    import torch
    cfg = _case("fixed_z")
    from peakpatch_hip.pipeline import TileDriver
    drv = TileDriver(cfg)
    z = torch.empty(0, dtype=torch.int32, device="cuda")
    pl = PeakList(z, z, z, z, torch.empty(0, dtype=torch.float32, device="cuda"), None)
    out = drv.device_batch(ctx, pl)
    assert all(len(t) == 0 for t in out.values())
