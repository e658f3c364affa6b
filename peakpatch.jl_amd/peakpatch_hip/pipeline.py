"""Tile drivers around the hot path, shaped like the reference's
``run_tile`` (src/Pipeline.jl:66-325) and the per-tile body of
``run_multitile_split`` (src/MultiResolution.jl:719-1180):

  filter bank + stencil (descending Rf, shared mask)  -> peak list
  per-peak Float64 host bookkeeping (ir2min, ZZon, fcrit, z_max cut)
  shell analysis (one batched call)                    -> PeakResult SoA
  record packing                                       -> HaloRecord / ExtHaloRecord

The workload configurations the benchmarks and tests use are defined here
(``CONFIGS``), each with its nbuff stated: nbuff = cap + 1 where cap =
floor(1.75 Rf_max / alatt) is the shell-radius bound of Pipeline.jl:96-98, so
every peak's shell sphere stays inside the tile.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import cosmology as cosmo_mod
from . import synth
from ._lib import SHELL_PERSISTENT


@dataclass
class TileConfig:
    name: str
    n: int
    boxsize: float
    nfilt: int
    rf_min: float
    rf_max: float
    ilpt: int = 2
    ioutshear: int = 1
    wsmooth: int = 1
    rmax2rs: float = 2.0
    ievol: int = 0
    z_out: float = 0.0
    z_max: float = 0.0
    cen: tuple = (0.0, 0.0, 0.0)          # observer position (light-cone mode)
    tile_center: tuple = (0.0, 0.0, 0.0)  # xbx, ybx, zbx
    seed: int = 42
    nbuff: int = 0                        # 0: cap + 1
    cosmology: cosmo_mod.CosmologyParams = field(default_factory=cosmo_mod.CosmologyParams)

    @property
    def alatt(self) -> float:
        return self.boxsize / self.n

    @property
    def filters(self) -> np.ndarray:
        return synth.filter_bank(self.nfilt, self.rf_min, self.rf_max)

    @property
    def cap(self) -> int:
        return int(math.floor(self.rf_max * 1.75 / self.alatt))

    @property
    def nbuff_eff(self) -> int:
        return self.nbuff if self.nbuff > 0 else self.cap + 1

    @property
    def nhunt(self) -> int:
        # Pipeline.jl:96-97
        return min(self.nbuff_eff - 1, int(math.floor(self.filters[0] * 1.75 / self.alatt)))

    def describe(self) -> str:
        mode = "lightcone (ievol=1)" if self.ievol == 1 else f"z={self.z_out}"
        return (f"{self.name}: box {self.boxsize:g} Mpc/h, nmesh={self.n}, {self.nfilt} top-hat filters "
                f"{self.rf_min:g}-{self.rf_max:g} Mpc/h, {self.ilpt}LPT, {mode}, ioutshear={self.ioutshear}, "
                f"rmax2rs={self.rmax2rs:g}, nbuff={self.nbuff_eff}, seed={self.seed}")


CONFIGS = {
    # BASELINE configs[0..3]
    "c1": TileConfig("c1", 128, 75.0, 10, 1.5, 8.0, ilpt=1, ioutshear=0, z_out=0.0),
    "c2": TileConfig("c2", 256, 250.0, 20, 1.0, 15.0, ilpt=2, ioutshear=1, z_out=0.5),
    "c3": TileConfig("c3", 512, 600.0, 25, 1.0, 30.0, ilpt=2, ioutshear=1, ievol=1, z_max=4.6,
                     cen=(-300.0, -300.0, -300.0)),
    "c4": TileConfig("c4", 768, 250.0, 20, 1.0, 15.0, ilpt=2, ioutshear=1, z_out=0.5),
}


def scaled(cfg: TileConfig, n: int) -> TileConfig:
    """Same physics at another mesh size (alatt changes; box fixed)."""
    return replace(cfg, n=n, name=f"{cfg.name}@{n}")


@dataclass
class PeakBatch:
    """Per-peak host arrays of one analyse_peaks call."""
    i: np.ndarray
    j: np.ndarray
    k: np.ndarray
    Rf: np.ndarray
    ir2min: np.ndarray
    ZZon: np.ndarray
    fcrit: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    FcRf: np.ndarray
    d2Rf: np.ndarray
    D: np.ndarray        # D_a at the peak (Dlinear_ab third output), 1LPT velocity factor
    v2fac: np.ndarray    # 3/7 Om_a^(-1/143) D^2, 2LPT velocity factor (zeros without 2LPT)

    def __len__(self):
        return len(self.i)

    def take(self, sel):
        return PeakBatch(*(getattr(self, f)[sel] for f in self.__dataclass_fields__))

    def to_device(self, device="cuda") -> dict:
        """The per-peak arrays analyse_peaks reads, as CUDA tensors."""
        import torch
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
        return {"i": t(self.i, torch.int32), "j": t(self.j, torch.int32), "k": t(self.k, torch.int32),
                "Rf": t(self.Rf, torch.float64), "ir2min": t(self.ir2min, torch.int64),
                "ZZon": t(self.ZZon, torch.float64), "fcrit": t(self.fcrit, torch.float64),
                "x": t(self.x, torch.float64), "y": t(self.y, torch.float64), "z": t(self.z, torch.float64),
                "FcRf": t(self.FcRf, torch.float32), "d2Rf": t(self.d2Rf, torch.float32),
                "D": t(self.D, torch.float64), "v2fac": t(self.v2fac, torch.float64)}


_TABLES: dict = {}


def _tables(cosmology, z_max: float, ievol: int):
    """Growth and chi->z tables depend on the cosmology only (built once per
    process, as the drivers build them once per run)."""
    key = (cosmology, z_max if ievol == 1 else None)
    t = _TABLES.get(key)
    if t is None:
        t = (cosmo_mod.GrowthTables(cosmology),
             cosmo_mod.ChiToZ(cosmology, z_max=z_max + 1.0) if ievol == 1 else None)
        _TABLES[key] = t
    return t


class TileDriver:
    """Holds the per-process host state a tile needs (growth tables, chi->z)."""

    def __init__(self, cfg: TileConfig):
        self.cfg = cfg
        self.growth, self.chi2z = _tables(cfg.cosmology, cfg.z_max, cfg.ievol)

    def filter_fcrit(self) -> np.ndarray:
        """fcrit per filter: fsc_of_z(z_out), or at the tile centre in light-cone
        mode (MultiResolution.jl:866-874)."""
        c = self.cfg
        if c.ievol == 1:
            zt = float(self.chi2z.peak_redshift(c.cen, *c.tile_center))
            f = float(self.growth.fsc_of_z(zt))
        else:
            f = float(self.growth.fsc_of_z(c.z_out))
        return np.full(c.nfilt, np.float32(f), np.float32)

    def find_peaks(self, ctx, delta, mask=None, device_out=False):
        c = self.cfg
        return ctx.find_peaks(delta, c.boxsize, c.filters, self.filter_fcrit(), wsmooth=c.wsmooth,
                              want_lapd=c.ioutshear >= 1, nbuff=c.nbuff_eff, mask=mask, device_out=device_out)

    def batch(self, peaks, round_f32: bool = False) -> PeakBatch:
        """Per-peak bookkeeping (Pipeline.jl:217-231, MultiResolution.jl:959-988).
        round_f32: ZZon and fcrit pass through Float32 as run_multitile_split
        stores them (MultiResolution.jl:963-967); run_tile keeps Float64."""
        c = self.cfg
        Rf = c.filters[peaks.filt]
        cen = 0.5 * (c.n + 1)
        x = c.tile_center[0] + c.alatt * ((peaks.i + 1) - cen)
        y = c.tile_center[1] + c.alatt * ((peaks.j + 1) - cen)
        z = c.tile_center[2] + c.alatt * ((peaks.k + 1) - cen)
        if c.ievol == 1:
            zpk = self.chi2z.peak_redshift(c.cen, x, y, z)
        else:
            zpk = np.full(len(x), c.z_out)
        a_pk = 1.0 / (1.0 + zpk)
        D_pk = self.growth.D_a(a_pk)
        if c.ilpt >= 2:
            # Pipeline.jl:268-271: -Sbar2 .* (-3/7 Om_a^(-1/143) D^2); the
            # two negations are exact, so the factor is kept positive
            Om, OL = c.cosmology.Om, c.cosmology.OL
            Om_a = Om * a_pk ** 3 / (Om * a_pk ** 3 + OL)
            v2 = -(-3.0 / 7.0 * Om_a ** (-1.0 / 143) * D_pk ** 2)
        else:
            v2 = np.zeros(len(x))
        ZZon, fcrit = 1.0 + zpk, self.growth.fsc_of_z(zpk)
        if round_f32:
            ZZon = ZZon.astype(np.float32).astype(np.float64)
            fcrit = fcrit.astype(np.float32).astype(np.float64)
        b = PeakBatch(peaks.i, peaks.j, peaks.k, Rf, cosmo_mod.ir2min(Rf, c.alatt), ZZon,
                      fcrit, x, y, z, peaks.delta_s,
                      peaks.lapd_s if peaks.lapd_s is not None else np.zeros(len(x), np.float32), D_pk, v2)
        if c.ievol == 1:
            b = b.take(zpk <= c.z_max)
        return b

    def geom(self, round_f32: bool = False):
        """The tile geometry and host tables pp_bookkeep_peaks reads; the
        numpy tables are kept alive on the driver."""
        from . import _lib as L
        c = self.cfg
        if getattr(self, "_geom", None) is None or self._geom[0] != round_f32:
            g = self.growth
            cols = [np.ascontiguousarray(a, np.float64) for a in (g.atab, g.Datab, g.dlnD)]
            chi = ([np.ascontiguousarray(self.chi2z.zs, np.float64), np.ascontiguousarray(self.chi2z.chis, np.float64)]
                   if c.ievol == 1 else [])
            gr = L.Growth(len(cols[0]), *(a.ctypes.data for a in cols))
            geo = L.TileGeom(c.n, c.alatt, *map(float, c.tile_center), int(c.ievol), float(c.z_out),
                             (C.c_double * 3)(*map(float, c.cen)), float(c.z_max), int(c.ilpt),
                             float(c.cosmology.Om), float(c.cosmology.OL), int(round_f32), gr,
                             len(chi[0]) if chi else 0, chi[0].ctypes.data if chi else None,
                             chi[1].ctypes.data if chi else None, float(self.chi2z.z_max) if chi else 0.0)
            self._geom = (round_f32, geo, cols, chi)
        return self._geom[1]

    def device_batch(self, ctx, peaks, round_f32: bool = False) -> dict:
        """batch() on the device: peaks is the device PeakList of
        find_peaks(..., device_out=True); returns PeakBatch.to_device()'s
        dict without a host round trip."""
        return ctx.bookkeep_peaks(peaks, self.cfg.filters, self.geom(round_f32))

    def analyse(self, ctx, fields, b, mask=None, kernel=SHELL_PERSISTENT, device_out=False):
        """b: a PeakBatch (host) or the dict of CUDA tensors of PeakBatch.to_device()."""
        c = self.cfg
        delta, eta, eta2, lapd = fields
        g = (lambda f: b[f]) if isinstance(b, dict) else (lambda f: getattr(b, f))
        return ctx.analyse_peaks(delta, eta, g("i"), g("j"), g("k"), g("Rf"), g("ir2min"), g("ZZon"), g("fcrit"),
                                 c.alatt, eta2=eta2 if c.ilpt >= 2 else None,
                                 lapd=lapd if c.ioutshear >= 1 else None, rmax2rs=c.rmax2rs, nbuff=c.nbuff_eff,
                                 mask=mask, kernel=kernel, device_out=device_out)

    def gather_cells(self, b: PeakBatch) -> np.ndarray:
        """Cells each peak's gather reads (centre + finalised shells), the unit
        of the algorithmic byte count."""
        return gather_cells(self.cfg.nhunt, b.Rf, self.cfg.alatt, self.cfg.rmax2rs)

    def gather_extent(self, b: PeakBatch) -> tuple[np.ndarray, np.ndarray]:
        return gather_extent(self.cfg.nhunt, b.Rf, self.cfg.alatt, self.cfg.rmax2rs)


def shell_counts(rmax: int) -> tuple[np.ndarray, np.ndarray]:
    """(r2 of each shell, cells per shell) of precompute_shells(rmax)."""
    r = np.arange(-rmax, rmax + 1)
    q = (r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2).ravel()
    q = q[q <= rmax * rmax]
    r2, cnt = np.unique(q, return_counts=True)
    return r2, cnt


def gather_cells(rmax: int, Rf, alatt: float, rmax2rs: float) -> np.ndarray:
    return gather_extent(rmax, Rf, alatt, rmax2rs)[0]


def gather_extent(rmax: int, Rf, alatt: float, rmax2rs: float) -> tuple[np.ndarray, np.ndarray]:
    """(cells read, finalised shells m) per peak: the centre plus shells
    2..m, m = shell(npart - 1) - 1 (RadialShell.jl:289-295, 379-453)."""
    r2, cnt = shell_counts(rmax)
    start = np.concatenate([[0], np.cumsum(cnt)])          # start[s-1] = first cell of shell s
    ncells = start[-1]
    Rf = np.asarray(Rf, np.float64)
    if rmax2rs > 0:
        t = rmax2rs * Rf / alatt
        npart = np.minimum(np.floor(4.0 / 3.0 * math.pi * (t * t * t)), ncells).astype(np.int64)
    else:
        npart = np.full(Rf.shape, ncells, np.int64)
    # m = shell(npart-1) - 1 finalised shells; cells read = start[m] (shells 1..m)
    shell_of_last = np.searchsorted(start, np.maximum(npart - 1, 0), side="right")
    m = np.where(npart >= 2, shell_of_last - 1, 1)
    return start[m], m
