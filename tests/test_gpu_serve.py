"""The tile-serving loop on the GPU: every tile's catalogue is the same
alone and in a mix of concurrent tiles (continuous batching over several
slots, work overlapping on the device), with the native server's own slot
contexts and its admission by device memory."""
import numpy as np
import pytest

from peakpatch_hip.serve import TileRequest, TileServer, replay, summarize

pytestmark = pytest.mark.gpu

TUNE = dict(small_m=12, chunk_shells=128, coop_m=256, gather_waves_cu=16)


def _requests():
    # light-cone tiles of uneven nmesh at the c5 resolution (nbuff = 27 for
    # periodic tiles, the split geometry's for split tiles)
    sizes = (96, 160, 128, 96, 192, 128)
    tiles = ((1, 1, 1), (2, 3, 1), (4, 1, 2), (3, 3, 3), (1, 4, 2), (2, 2, 4))
    return [TileRequest(q, n, center=(300.0 + 90 * q, 200.0, 150.0 + 40 * q), seed=77 + q, tile=t)
            for q, (n, t) in enumerate(zip(sizes, tiles))]


@pytest.mark.parametrize("split", [False, True], ids=["periodic", "split"])
def test_tiles_same_alone_and_in_mix(homeltab_path, split):
    from peakpatch_hip import collapse
    from peakpatch_hip.serve import GpuTileRunner
    runner = GpuTileRunner(table=collapse.read_homeltab(homeltab_path), tunables=TUNE, split=split)
    # the server's own slot contexts, admission by device memory (budget:
    # a fraction of the free HBM, and the live free memory per tile)
    # alone: one slot, each tile submitted after the previous one finished
    with TileServer(slots=1, runner=runner) as srv:
        assert srv.budget > 0
        alone = [srv.submit(r).result(timeout=300) for r in _requests()]
    assert all(r.error is None for r in alone), [r.error for r in alone]
    # mix: three slots, all tiles arriving together
    with TileServer(slots=3, runner=runner) as srv:
        mix = replay(srv, [(0.0, r) for r in _requests()])
        overlap = srv.max_in_flight
        st = srv.stats()
    assert overlap == 3 and st["finished"] == 6 and st["failed"] == 0
    assert all(r.error is None for r in alone + mix), [r.error for r in alone + mix]
    assert sum(r.nhalo for r in alone) > 50
    for a, m in zip(alone, mix):
        assert a.tile_id == m.tile_id and a.npeaks == m.npeaks and a.nhalo == m.nhalo
        # (gradpkrf is NaN where the reference divides by rad[1] = 0)
        assert np.array_equal(a.records, m.records, equal_nan=True), f"tile {a.tile_id} differs in the mix"
        assert a.RTHLmax == m.RTHLmax
    s = summarize(mix)
    assert s["completed"] == 6 and s["p99_ms"] >= s["p50_ms"] > 0
    runner.close()
