"""Multi-resolution tile fields: run_multitile_split's split of one global
Threefry field into a device-resident coarse grid and per-tile isolated
convolutions (src/MultiResolution.jl:523-850), on the library's
pp_coarse_* / pp_split_tile_fields entry points.

A ``SplitGeom`` fixes the global grid (N = nsub ntile + 2 nbuff, coarse M
dividing N); ``CoarseField`` holds its coarse fields on one device for the
life of a run (or a server); ``make_split_tile`` generates one tile's fields
into CUDA tensors.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

from . import _lib as L


@dataclass(frozen=True)
class SplitGeom:
    nmesh: int
    nbuff: int
    ntile: int
    M: int
    alatt: float
    seed: int = 42

    @property
    def nsub(self) -> int:
        return self.nmesh - 2 * self.nbuff

    @property
    def N(self) -> int:
        return self.nsub * self.ntile + 2 * self.nbuff

    @property
    def block(self) -> int:
        return self.N // self.M

    @property
    def boxsize_local(self) -> float:
        return self.nmesh * self.alatt

    @property
    def dcore_box(self) -> float:
        return self.nsub * self.alatt

    def tile_center(self, it: int, jt: int, kt: int):
        """tile_center (src/MultiTile.jl:79-86), tiles 1-based."""
        cen = (self.ntile + 1) / 2
        return tuple((t - cen) * self.dcore_box for t in (it, jt, kt))

    def validate(self):
        if self.nsub < 1 or self.M < 2 or self.N % self.M:
            raise ValueError(f"N={self.N} must be divisible by M={self.M} (nsub={self.nsub})")

    def c_struct(self) -> L.SplitGeom:
        return L.SplitGeom(self.nmesh, self.nbuff, self.ntile, self.M, float(self.alatt),
                           int(self.seed) & 0xFFFFFFFFFFFFFFFF)


def geometry(nmesh: int, nbuff_min: int, ntile: int, alatt: float, coarse_factor: int = 5,
             seed: int = 42) -> SplitGeom:
    """run_multitile_split's geometry with M = coarse_factor x ntile
    (MultiResolution.jl:596-599): the smallest nbuff >= nbuff_min for which
    M divides N."""
    M = coarse_factor * ntile
    for nbuff in range(nbuff_min, nmesh // 2):
        g = SplitGeom(nmesh, nbuff, ntile, M, alatt, seed)
        if g.nsub >= 1 and g.N % M == 0:
            return g
    raise ValueError(f"no nbuff >= {nbuff_min} makes N divisible by M={M} for nmesh={nmesh}, ntile={ntile}")


class CoarseField:
    """The coarse noise, delta_coarse and psi_coarse of a split geometry on
    ctx's device (phase 1a, MultiResolution.jl:617-631)."""

    def __init__(self, ctx, geom: SplitGeom, pk=None):
        from . import synth
        geom.validate()
        self.geom = geom
        self.device = ctx.device
        self._lib = ctx.lib
        k, P = synth.pk_table(pk)
        tab = L.PkTable(len(k), k.ctypes.data, P.ctypes.data)
        h = C.c_void_p()
        L.check(ctx.lib.pp_coarse_create(ctx.h, C.byref(geom.c_struct()), C.byref(tab), C.byref(h)))
        self.h = h

    def download(self):
        """(noise, delta, (psi_x, psi_y, psi_z)) as host (M, M, M) float32 arrays [k, j, i]."""
        import numpy as np
        M = self.geom.M
        a = [np.empty((M, M, M), np.float32) for _ in range(5)]
        L.check(self._lib.pp_coarse_download(self.h, *(x.ctypes.data for x in a)))
        return a[0], a[1], tuple(a[2:])

    def close(self):
        if getattr(self, "h", None):
            self._lib.pp_coarse_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_split_tile(ctx, coarse: CoarseField, it: int, jt: int, kt: int, ilpt: int = 2, want_lapd: bool = True):
    """One tile's fields (process_tile!, MultiResolution.jl:719-850) as CUDA
    float32 tensors on ctx's device, ordered before later work on torch's
    current stream."""
    import torch

    from .synth import TileInputs
    g = coarse.geom
    n = g.nmesh
    dev = torch.device("cuda", ctx.device)
    new = lambda: torch.empty((n, n, n), dtype=torch.float32, device=dev)
    delta = new()
    eta = (new(), new(), new())
    eta2 = (new(), new(), new()) if ilpt >= 2 else None
    lapd = new() if want_lapd else None
    p = lambda t: t.data_ptr() if t is not None else None
    f = L.TileFields(p(delta), *(p(a) for a in eta), *((p(a) for a in eta2) if eta2 else (None, None, None)), p(lapd))
    ctx._call(lambda: ctx.lib.pp_split_tile_fields(ctx.h, coarse.h, int(it), int(jt), int(kt), C.byref(f)), True)
    return TileInputs(n, g.boxsize_local, delta, eta, eta2, lapd)


def footprint_bytes(n: int) -> int:
    """Device bytes a split tile of nmesh n holds at its peak: two padded
    (2n)^3 spectra (64 n^3 B each), the eight output fields (32 n^3), the
    tile-periodic spectra and real scratch (16 n^3), then the peak finder's
    and the shell analysis' working set (~60 n^3 at ~1.5e-2 peaks per cell),
    plus fixed tables and plans."""
    return int(n) ** 3 * 172 + (512 << 20)


def tiles(geom: SplitGeom):
    """The run's tile indices in run_multitile_split's order (kt, jt, it
    nested, it fastest; MultiResolution.jl:672-676)."""
    return [(it, jt, kt) for kt in range(1, geom.ntile + 1) for jt in range(1, geom.ntile + 1)
            for it in range(1, geom.ntile + 1)]


def tile_distance(geom: SplitGeom, t, obs) -> float:
    """Distance from obs to the nearest point of tile t's core
    (MultiResolution.jl:677-687)."""
    half = geom.dcore_box / 2.0
    c = geom.tile_center(*t)
    d = [max(abs(c[q] - obs[q]) - half, 0.0) for q in range(3)]
    return math.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
