"""GPU record packing (pp_pack_halos) against the oracle's restatement of
the drivers' packing loop: bit-exact Float32 records."""
import numpy as np
import pytest

from conftest import ipp1
from oracle import catalog as OC
from oracle import oracle as O
from peakpatch_hip import PPError, SHELL_PERSISTENT, catalog, collapse
from test_gpu_shell import bump_field, upload_table

pytestmark = pytest.mark.gpu


def _case(ctx):
    import torch
    n, rmax = 48, 12
    d, eta, eta2, lapd = bump_field(n, 17, nbump=25, noise=0.3)
    upload_table(ctx, collapse.sinusoidal_table())
    ctx.build_shell_tables(rmax)
    i, j, k, _ = O.find_peaks(d, np.zeros(d.shape, np.int8), 1, 0.6)
    i, j, k = i - 1, j - 1, k - 1
    rng = np.random.default_rng(2)
    npk = len(i)
    Rf = rng.choice([1.0, 1.7, 2.9, 4.4], npk)
    ir2 = np.floor((1.75 * Rf) ** 2).astype(np.int64)
    ZZ = 0.3 + 1.2 * rng.random(npk)
    fc = 0.6 + 0.3 * rng.random(npk)
    cud = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")
    host = {"x": rng.uniform(-100, 100, npk), "y": rng.uniform(-100, 100, npk), "z": rng.uniform(-100, 100, npk),
            "Rf": Rf, "ZZon": ZZ, "D": rng.uniform(0.5, 1.0, npk), "v2fac": rng.uniform(0.3, 0.5, npk),
            "FcRf": rng.normal(size=npk).astype(np.float32), "d2Rf": rng.normal(size=npk).astype(np.float32)}
    bdev = {kk: cud(v, torch.float32 if v.dtype == np.float32 else torch.float64) for kk, v in host.items()}
    f = lambda a: cud(a, torch.float32)
    res = ctx.analyse_peaks(f(d), tuple(f(a) for a in eta), cud(i, torch.int32), cud(j, torch.int32),
                            cud(k, torch.int32), bdev["Rf"], cud(ir2, torch.int64), bdev["ZZon"], cud(fc, torch.float64),
                            1.0, eta2=tuple(f(a) for a in eta2), lapd=f(lapd), rmax2rs=2.0, kernel=SHELL_PERSISTENT,
                            device_out=True)
    return res, bdev, host


@pytest.mark.parametrize("extended", [False, True])
def test_pack_halos_bitexact(ctx, extended):
    res, bdev, host = _case(ctx)
    rec, rmax = catalog.pack_halos(ctx, res, bdev, 1.17, extended)
    hres = {kk: (v.cpu().numpy() if hasattr(v, "cpu") else v) for kk, v in res.items()}
    ref, rref = OC.pack(hres, host["x"], host["y"], host["z"], host["Rf"], host["ZZon"], host["D"], host["v2fac"],
                        host["FcRf"], host["d2Rf"], 1.17, extended)
    assert len(ref) > 20
    # gradpkrf is NaN where the reference divides by rad[1] = 0 (NaN == NaN here)
    assert np.array_equal(rec.cpu().numpy(), ref, equal_nan=True)
    assert rmax == rref


def test_pack_halos_zmax_cut_and_capacity(ctx):
    res, bdev, host = _case(ctx)
    hres = {kk: (v.cpu().numpy() if hasattr(v, "cpu") else v) for kk, v in res.items()}
    zcut = float(np.median(host["ZZon"][hres["RTHL"] > 0]) - 1.0)
    rec, rmax = catalog.pack_halos(ctx, res, bdev, 1.0, False, z_max=zcut)
    ref, rref = OC.pack(hres, host["x"], host["y"], host["z"], host["Rf"], host["ZZon"], host["D"], host["v2fac"],
                        host["FcRf"], host["d2Rf"], 1.0, False, z_max=zcut)
    assert 0 < len(ref) < int((hres["RTHL"] > 0).sum())
    assert np.array_equal(rec.cpu().numpy(), ref, equal_nan=True) and rmax == rref
    # a buffer too small is an error, not a truncation
    import ctypes as C
    from peakpatch_hip import _lib as L
    import torch
    p = lambda t: t.data_ptr()
    r = L.Results(*(p(res[f]) for f in catalog._RESULT_ORDER), p(res["status"]), p(res["nshells"]), 1)
    h = L.HaloInputs(len(host["x"]), *(p(bdev[kk]) for kk in ("x", "y", "z", "Rf", "ZZon", "D", "v2fac", "FcRf",
                                                                 "d2Rf")), 1)
    out = torch.empty((4, 11), dtype=torch.float32, device="cuda")
    nh, rm = C.c_int64(), C.c_float()
    rc = ctx.lib.pp_pack_halos(ctx.h, C.byref(r), C.byref(h), 1.0, float("inf"), 0, p(out), 4, 1, C.byref(nh),
                               C.byref(rm))
    assert rc == -6 and nh.value == int((hres["RTHL"] > 0).sum())
