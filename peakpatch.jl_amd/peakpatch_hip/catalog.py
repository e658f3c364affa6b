"""Halo catalogue of the path: record packing on the device (pp_pack_halos,
src/Pipeline.jl:247-308 and src/MultiResolution.jl:1025-1094) and the .pksc
file format (src/IO/Catalog.jl:48-160), byte-compatible with write_pksc /
read_pksc: Int32 Nhalo, Float32 RTHLmax, Float32 z_out, then Nhalo records
of 11 (HaloRecord) or 33 (ExtHaloRecord) Float32, little-endian.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L

BASIC_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "RTHL", "vx2", "vy2", "vz2", "overdensity")
EXT_FIELDS = BASIC_FIELDS + ("e_v", "p_v", "strain_11", "strain_22", "strain_33", "strain_23", "strain_13",
                             "strain_12", "d2F", "zform", "grad_x", "grad_y", "grad_z", "gradf_x", "gradf_y",
                             "gradf_z", "Rf", "FcollvRf", "d2FRf", "gradrf_x", "gradrf_y", "gradrf_z")

_RESULT_ORDER = ("RTHL", "Srb", "Sbar", "Sbar2", "strain", "gradpk", "gradpkrf", "Fbarx", "e_v", "p_v",
                 "zvir_half", "gradpkf", "d2F")


def pack_halos(ctx, res: dict, bdev: dict, alatt: float, extended: bool, z_max: float = math.inf):
    """Records of the collapsed peaks, in peak order, as a CUDA tensor
    (nhalo, 11 | 33) plus RTHLmax.  res: the device results of
    Context.analyse_peaks(..., device_out=True); bdev: PeakBatch.to_device();
    peaks with z > z_max are dropped (the light-cone cut)."""
    import torch
    n = int(bdev["x"].numel())
    p = lambda t: t.data_ptr()
    r = L.Results(*(p(res[f]) for f in _RESULT_ORDER), p(res["status"]), p(res["nshells"]), 1)
    h = L.HaloInputs(n, p(bdev["x"]), p(bdev["y"]), p(bdev["z"]), p(bdev["Rf"]), p(bdev["ZZon"]), p(bdev["D"]),
                     p(bdev["v2fac"]) if "v2fac" in bdev and bdev["v2fac"] is not None else None,
                     p(bdev["FcRf"]), p(bdev["d2Rf"]) if bdev.get("d2Rf") is not None else None, 1)
    nf = 33 if extended else 11
    nh = C.c_int64()
    rmax = C.c_float()
    # size the output by the collapsed count (the pack call reports it)
    cap = int((res["RTHL"] > 0).sum().item())
    out = torch.empty((max(cap, 1), nf), dtype=torch.float32, device=bdev["x"].device)
    ctx._call(lambda: ctx.lib.pp_pack_halos(ctx.h, C.byref(r), C.byref(h), alatt, z_max, 1 if extended else 0,
                                            out.data_ptr(), cap, 1, C.byref(nh), C.byref(rmax)), True)
    return out[: nh.value], float(rmax.value)


def merge_catalog(ctx, records):
    """merge_catalog (src/Merger/Merger.jl:21-79) on the device
    (pp_merge_halos): Lagrangian exclusion and volume reduction of a
    catalogue of HaloRecord / ExtHaloRecord rows (numpy (nhalo, 11 | 33) or a
    CUDA tensor).  Returns the surviving rows in catalogue order with RTHL
    replaced by the reduced radius, of the same kind as the input."""
    on_dev = type(records).__module__.startswith("torch")
    n = int(records.shape[0])
    if n == 0:
        return records[:0]
    nk = C.c_int64()
    if on_dev:
        import torch
        cols = [records[:, q].contiguous() for q in (0, 1, 2, 6)]
        keep = torch.empty(n, dtype=torch.int8, device=records.device)
        nr = torch.empty(n, dtype=torch.float32, device=records.device)
        ctx._call(lambda: ctx.lib.pp_merge_halos(ctx.h, n, *(c.data_ptr() for c in cols), 1, keep.data_ptr(),
                                                 nr.data_ptr(), C.byref(nk)), True)
        sel = keep.bool()
        out = records[sel].clone()
        out[:, 6] = nr[sel]
        return out
    rec = np.ascontiguousarray(records, np.float32)
    cols = [np.ascontiguousarray(rec[:, q]) for q in (0, 1, 2, 6)]
    keep = np.zeros(n, np.int8)
    nr = np.zeros(n, np.float32)
    L.check(ctx.lib.pp_merge_halos(ctx.h, n, *(c.ctypes.data for c in cols), 0, keep.ctypes.data, nr.ctypes.data,
                                   C.byref(nk)))
    sel = keep.astype(bool)
    out = rec[sel].copy()
    out[:, 6] = nr[sel]
    return out


def write_pksc(path: str, records: np.ndarray, RTHLmax: float, z_out: float) -> None:
    """write_pksc (Catalog.jl:53-94)."""
    rec = np.ascontiguousarray(records, dtype="<f4")
    if rec.ndim != 2 or rec.shape[1] not in (11, 33):
        raise ValueError("records must be (Nhalo, 11) or (Nhalo, 33)")
    with open(path, "wb") as f:
        f.write(np.array([rec.shape[0]], "<i4").tobytes())
        f.write(np.array([RTHLmax, z_out], "<f4").tobytes())
        f.write(rec.tobytes())


def read_pksc(path: str):
    """read_pksc (Catalog.jl:100-160): (records, RTHLmax, z_out); the record
    width comes from the file size (33 or else 11)."""
    raw = open(path, "rb").read()
    nh = int(np.frombuffer(raw[:4], "<i4")[0])
    RTHLmax, z_out = (float(v) for v in np.frombuffer(raw[4:12], "<f4"))
    nf = (len(raw) - 12) // (nh * 4) if nh > 0 else 11
    nf = 33 if nf == 33 else 11
    rec = np.frombuffer(raw[12:12 + nh * nf * 4], "<f4").reshape(nh, nf).copy()
    return rec, np.float32(RTHLmax), np.float32(z_out)


def halo_mass(RTHL_phys, Om: float, h: float):
    """M200m of the HDF5 catalogue (ext/HDF5Ext.jl:58, 68):
    4/3 pi (2.775e11 Om h^2) RTHL^3, RTHL in Mpc/h (the h^2 is the
    reference's, quirk Q15)."""
    rho = 2.775e11 * Om * h * h
    return 4.0 / 3.0 * math.pi * rho * np.asarray(RTHL_phys, np.float64) ** 3
