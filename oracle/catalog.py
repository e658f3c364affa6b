"""TEST INFRASTRUCTURE ONLY (the oracle): record packing of the drivers
restated in numpy, for parity tests of pp_pack_halos.  Follows
src/Pipeline.jl:247-308 (run_tile) and src/MultiResolution.jl:1025-1094
literally: per collapsed peak, Float64 products cast to Float32 in the
HaloRecord / ExtHaloRecord field order of src/IO/Catalog.jl:5-46.
"""
import numpy as np


def pack(res: dict, x, y, z, Rf, ZZon, D, v2fac, FcRf, d2Rf, alatt: float, extended: bool, z_max: float = np.inf):
    """res: PeakResult arrays keyed as Context.analyse_peaks returns them
    (strain as (n, 3, 3) strain_mat).  Returns (records (nh, 11|33) float32,
    RTHLmax float32)."""
    rows = []
    for p in range(len(x)):
        if not res["RTHL"][p] > 0:                      # Pipeline.jl:249
            continue
        if ZZon[p] - 1.0 > z_max:                       # Pipeline.jl:252-255
            continue
        f32 = np.float32
        RTHL_phys = f32(res["RTHL"][p] * alatt)
        vel = [f32(res["Sbar"][p][L] * D[p]) for L in range(3)]
        vel2 = [f32(res["Sbar2"][p][L] * v2fac[p]) if v2fac is not None else f32(0) for L in range(3)]
        r = [f32(x[p]), f32(y[p]), f32(z[p]), *vel, RTHL_phys, *vel2, f32(res["Fbarx"][p])]
        if extended:
            sm = res["strain"][p]
            r += [f32(res["e_v"][p]), f32(res["p_v"][p]),
                  f32(sm[0, 0]), f32(sm[1, 1]), f32(sm[2, 2]), f32(sm[1, 2]), f32(sm[0, 2]), f32(sm[0, 1]),
                  f32(res["d2F"][p]), f32(res["zvir_half"][p]),
                  *[f32(res["gradpk"][p][L]) for L in range(3)], *[f32(res["gradpkf"][p][L]) for L in range(3)],
                  f32(Rf[p]), f32(FcRf[p]), f32(d2Rf[p]) if d2Rf is not None else f32(0),
                  *[f32(res["gradpkrf"][p][L]) for L in range(3)]]
        rows.append(r)
    nf = 33 if extended else 11
    rec = np.array(rows, np.float32).reshape(-1, nf)
    rmax = rec[:, 6].max() if len(rec) else np.float32(0)  # Pipeline.jl:308-313
    return rec, np.float32(rmax)
