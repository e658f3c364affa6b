"""Multi-resolution tile fields on the GPU (pp_coarse_create,
pp_split_tile_fields) against the oracle's restatement of
run_multitile_split (src/MultiResolution.jl:41-314, 523-850).

Tolerances: the coarse noise is a Float64 sum rounded to Float32 (the
oracle's sum associates differently), so it agrees to one Float32 ulp; every
other field passes through Float32 transforms (rocFFT here, pocketfft in
Float64 in the oracle), so it agrees to 2e-5 of the field's largest value.
Two runs on the device are bit-identical.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from peakpatch_hip.api import Context
    return Context(0)


def _pk():
    from oracle import fields as OF
    from peakpatch_hip import synth
    return OF.PkTable(*synth.pk_table())


GEOMS = {
    # nmesh, nbuff, ntile, M, alatt: block 8 (coarse_factor 5); block 5 (odd) with a middle tile;
    # a coarse grid too large for LDS (32^3 floats: the interpolation reads it from global memory)
    "cf5": (48, 8, 2, 10, 2.0),
    "odd": (40, 5, 3, 20, 1.5),
    "big": (32, 0, 2, 32, 2.0),
}


def _close(name, got, want, rel=2e-5):
    scale = float(np.abs(want).max())
    err = float(np.abs(got - want).max())
    assert scale > 0 and err <= rel * scale, (name, err, scale)


@pytest.mark.parametrize("gname", sorted(GEOMS))
def test_coarse_fields_match_the_oracle(ctx, gname):
    from oracle import fields as OF
    from peakpatch_hip import split
    nmesh, nbuff, ntile, M, alatt = GEOMS[gname]
    g = split.SplitGeom(nmesh, nbuff, ntile, M, alatt, seed=21)
    cf = split.CoarseField(ctx, g)
    noise, delta, psi = cf.download()
    cf.close()
    og = OF.SplitGeom(nmesh, nbuff, ntile, M, alatt)
    on, od, op = OF.coarse_fields(og, 21, _pk())
    print(f"\n[split] {gname}: N={g.N} M={M} block={g.block}", flush=True)
    ulp = np.spacing(np.abs(on))
    assert np.all(np.abs(noise - on) <= ulp), float(np.abs(noise - on).max())
    _close("delta_coarse", delta, od)
    for q in range(3):
        _close(f"psi_coarse{q}", psi[q], op[q])


@pytest.mark.parametrize("gname,tile", [("cf5", (1, 1, 1)), ("cf5", (2, 1, 2)), ("odd", (2, 3, 1)),
                                        ("odd", (3, 2, 2)), ("big", (2, 1, 2))])
def test_tile_fields_match_the_oracle(ctx, gname, tile):
    import torch
    from oracle import fields as OF
    from peakpatch_hip import split
    nmesh, nbuff, ntile, M, alatt = GEOMS[gname]
    g = split.SplitGeom(nmesh, nbuff, ntile, M, alatt, seed=33)
    cf = split.CoarseField(ctx, g)
    t1 = split.make_split_tile(ctx, cf, *tile)
    t2 = split.make_split_tile(ctx, cf, *tile)
    torch.cuda.synchronize()
    og = OF.SplitGeom(nmesh, nbuff, ntile, M, alatt)
    ref = OF.split_tile(og, *tile, 33, _pk())
    pairs = [("delta", t1.delta, t2.delta, ref.delta), ("lapd", t1.lapd, t2.lapd, ref.lapd)]
    pairs += [(f"eta{q}", a, b, c) for q, (a, b, c) in enumerate(zip(t1.eta, t2.eta, ref.eta))]
    pairs += [(f"eta2{q}", a, b, c) for q, (a, b, c) in enumerate(zip(t1.eta2, t2.eta2, ref.eta2))]
    for name, a, b, c in pairs:
        assert torch.equal(a, b), f"{name} differs between two runs"
        _close(name, a.cpu().numpy(), c)
    cf.close()


def test_serving_sized_tile(ctx):
    """A c5-sized tile (nmesh 256 at alatt 250/256, coarse_factor 5): finite
    fields, the same twice, and the coarse part carrying the long modes (the
    tile mean of delta is the interpolated coarse field's, not zero)."""
    import torch
    from peakpatch_hip import split
    g = split.geometry(256, 27, 2, 250.0 / 256, coarse_factor=5, seed=5)
    cf = split.CoarseField(ctx, g)
    a = split.make_split_tile(ctx, cf, 2, 1, 2)
    b = split.make_split_tile(ctx, cf, 2, 1, 2)
    torch.cuda.synchronize()
    for x, y in [(a.delta, b.delta), (a.lapd, b.lapd), *zip(a.eta, b.eta), *zip(a.eta2, b.eta2)]:
        assert torch.equal(x, y) and bool(torch.isfinite(x).all())
    assert float(a.delta.std()) > 0.1
    cf.close()


def test_bad_tile_index_is_refused(ctx):
    from peakpatch_hip import PPError, split
    g = split.SplitGeom(*GEOMS["cf5"], seed=1)
    cf = split.CoarseField(ctx, g)
    with pytest.raises(PPError):
        split.make_split_tile(ctx, cf, 3, 1, 1)
    cf.close()


def test_release_scratch_keeps_results(ctx):
    """pp_ctx_release_scratch returns the pooled memory; the next calls grow
    the pools again and give the same fields and peaks."""
    import torch
    from peakpatch_hip import split
    from peakpatch_hip.pipeline import CONFIGS, TileDriver, scaled
    g = split.SplitGeom(*GEOMS["cf5"], seed=3)
    cf = split.CoarseField(ctx, g)
    cfg = scaled(CONFIGS["c2"], g.nmesh)
    a = split.make_split_tile(ctx, cf, 1, 2, 1)
    pa = TileDriver(cfg).find_peaks(ctx, a.delta)
    ctx.release_scratch()
    b = split.make_split_tile(ctx, cf, 1, 2, 1)
    pb = TileDriver(cfg).find_peaks(ctx, b.delta)
    torch.cuda.synchronize()
    assert torch.equal(a.delta, b.delta) and torch.equal(a.eta2[1], b.eta2[1])
    assert len(pa) > 0 and np.array_equal(pa.i, pb.i) and np.array_equal(pa.delta_s, pb.delta_s)
    cf.close()
