"""The catalogue merger on the GPU (pp_merge_halos) against the oracle's
restatement of merge_catalog (src/Merger/Merger.jl:21-79,
src/Merger/Exclusion.jl).

The surviving set must be identical; the reduced radius, Float64 rounded to
Float32 as _replace_rthl stores it, must agree to one Float32 ulp (the
device sums each halo's caps in another order than the reference's serial
sweep).  Two runs are bit-identical, and host and device inputs give the
same catalogue.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from peakpatch_hip.api import Context
    return Context(0)


def _catalogue(seed, n, box=200.0, spread=4.0, nclust=12):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, box, (nclust, 3))
    pos = centres[rng.integers(0, nclust, n)] + rng.normal(0, spread, (n, 3))
    rec = np.zeros((n, 11), np.float32)
    rec[:, :3] = pos
    rec[:, 6] = np.minimum(rng.pareto(2.5, n) + 1.0, 3.0) * 1.2
    rec[:5, 6] = rec[5, 6]                  # exact ties
    rec[:, 3:6] = rng.normal(0, 100, (n, 3))
    rec[:, 10] = rng.uniform(1, 3, n)
    return rec


def _check(rec, got):
    from oracle import merger as OM
    alive, nr = OM.merge(rec[:, 0], rec[:, 1], rec[:, 2], rec[:, 6])
    want = rec[alive].copy()
    want[:, 6] = nr[alive].astype(np.float32)
    assert got.shape == want.shape, (got.shape, want.shape)
    other = [q for q in range(rec.shape[1]) if q != 6]
    np.testing.assert_array_equal(got[:, other], want[:, other])
    ulp = np.spacing(want[:, 6])
    assert np.all(np.abs(got[:, 6] - want[:, 6]) <= ulp), float(np.abs(got[:, 6] - want[:, 6]).max())
    return int(alive.sum())


@pytest.mark.parametrize("seed,n", [(1, 300), (2, 400), (3, 1500)])
def test_merge_matches_the_oracle(ctx, seed, n):
    from peakpatch_hip import catalog
    rec = _catalogue(seed, n)
    got = catalog.merge_catalog(ctx, rec)
    kept = _check(rec, got)
    print(f"\n[merge] seed {seed}: {n} -> {kept} halos", flush=True)
    assert 0 < kept < n
    again = catalog.merge_catalog(ctx, rec)
    np.testing.assert_array_equal(got, again)


def test_merge_device_input(ctx):
    import torch
    from peakpatch_hip import catalog
    rec = _catalogue(7, 800)
    host = catalog.merge_catalog(ctx, rec)
    dev = catalog.merge_catalog(ctx, torch.as_tensor(rec, device="cuda"))
    np.testing.assert_array_equal(dev.cpu().numpy(), host)


def test_merge_tile_catalogue(ctx, homeltab_path):
    """A catalogue from the pipeline (a 128^3 tile at the c2 resolution,
    extended records), a sub-box of it through the oracle."""
    import torch
    from peakpatch_hip import catalog, collapse, synth
    from peakpatch_hip.pipeline import CONFIGS, TileDriver, scaled
    cfg = scaled(CONFIGS["c2"], 128)
    ct = collapse.read_homeltab(homeltab_path)
    ctx.upload_collapse_table(ct.table, *ct.bounds)
    ctx.build_shell_tables(cfg.nhunt)
    t = synth.make_tile(ctx, cfg.n, cfg.boxsize, cfg.seed, ilpt=cfg.ilpt, want_lapd=True)
    drv = TileDriver(cfg)
    peaks = drv.find_peaks(ctx, t.delta, device_out=True)
    b = drv.device_batch(ctx, peaks)
    mask = torch.zeros((cfg.n,) * 3, dtype=torch.int8, device="cuda")
    res = drv.analyse(ctx, (t.delta, t.eta, t.eta2, t.lapd), b, mask=mask, device_out=True)
    rec, _ = catalog.pack_halos(ctx, res, b, cfg.alatt, extended=True)
    rec = rec.cpu().numpy()
    sub = rec[(np.abs(rec[:, :3]) < 45.0).all(axis=1)]
    print(f"\n[merge] tile catalogue: {len(rec)} halos, sub-box {len(sub)}", flush=True)
    assert len(sub) > 200
    got = catalog.merge_catalog(ctx, sub)
    kept = _check(sub, got)
    assert 0 < kept < len(sub)
    full = catalog.merge_catalog(ctx, rec)
    assert 0 < len(full) < len(rec)


def test_merge_empty(ctx):
    from peakpatch_hip import catalog
    out = catalog.merge_catalog(ctx, np.zeros((0, 11), np.float32))
    assert out.shape == (0, 11)
