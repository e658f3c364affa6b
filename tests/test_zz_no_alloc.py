"""pp_find_peaks_field_noalloc -- the peak stencil with its scratch in a
static device array, no allocation -- against pp_find_peaks_field on the same
inputs: two calls in a row, the second on a smaller grid, then a grid beyond
the scratch limit, which is refused before any work is queued."""
import ctypes as C

import numpy as np
import pytest

from peakpatch_hip import Context
from peakpatch_hip import _lib as L

pytestmark = pytest.mark.gpu


def _field(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, n, n)).astype(np.float32)
    # a little smoothing so that peaks are a few per cent of the cells
    s = d.copy()
    for ax in range(3):
        s += np.roll(d, 1, axis=ax) + np.roll(d, -1, axis=ax)
    return (s / 7.0).astype(np.float32)


def _noalloc(ctx, d, mask, nbuff, fcrit, cap):
    import torch
    n3, n2, n1 = d.shape
    dd = torch.from_numpy(d).cuda()
    md = torch.from_numpy(mask).cuda()
    out = {q: torch.full((cap,), -7, dtype=torch.int32, device="cuda") for q in ("i", "j", "k", "filt")}
    out["delta_s"] = torch.full((cap,), np.nan, dtype=torch.float32, device="cuda")
    pk = L.Peaks(cap, 0, 0, *(out[q].data_ptr() for q in ("i", "j", "k", "filt", "delta_s")), None)
    torch.cuda.synchronize()
    L.check(ctx.lib.pp_find_peaks_field_noalloc(ctx.h, dd.data_ptr(), md.data_ptr(), n1, n2, n3, nbuff,
                                                C.c_float(fcrit), 2_000_000, C.byref(pk)))
    host = {q: v.cpu().numpy()[: pk.count] for q, v in out.items()}
    return pk, host, md.cpu().numpy()


def test_noalloc_matches_the_allocating_entry_point():
    with Context(0) as ctx:
        for n, nbuff, seed in ((96, 3, 11), (64, 2, 12)):
            d = _field(n, seed)
            m0 = np.zeros(d.shape, np.int8)
            m0[::7, ::5, ::3] = 1                  # some cells masked by earlier filters
            m_ref = m0.copy()
            ref = ctx.find_peaks_field(d, m_ref, nbuff, 0.4)
            assert len(ref) > 100
            pk, got, m_got = _noalloc(ctx, d, m0.copy(), nbuff, 0.4, len(ref) + 16)
            assert pk.count == pk.needed == len(ref)
            assert np.array_equal(got["i"], ref.i) and np.array_equal(got["j"], ref.j)
            assert np.array_equal(got["k"], ref.k) and np.array_equal(got["filt"], ref.filt)
            assert np.array_equal(got["delta_s"], ref.delta_s)
            assert np.array_equal(m_got, m_ref)


def test_noalloc_refuses_grids_beyond_its_scratch():
    import torch
    with Context(0) as ctx:
        tiny = torch.zeros(8, dtype=torch.float32, device="cuda")
        m = torch.zeros(8, dtype=torch.int8, device="cuda")
        o = torch.zeros(8, dtype=torch.int32, device="cuda")
        pk = L.Peaks(8, 0, 0, o.data_ptr(), o.data_ptr(), o.data_ptr(), o.data_ptr(), tiny.data_ptr(), None)
        rc = ctx.lib.pp_find_peaks_field_noalloc(ctx.h, tiny.data_ptr(), m.data_ptr(), 1600, 1600, 1600, 2,
                                                 C.c_float(0.0), 0, C.byref(pk))
        assert rc == -1 and b"scratch" in ctx.lib.pp_last_error()
