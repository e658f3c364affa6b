"""Seeded synthetic inputs for the halo-finding path.

The reference reads P(k) from a two-column file (load_pk,
src/InitialConditions/PowerSpectrum.jl:11-23) and builds a tile's fields in
run_tile (src/Pipeline.jl:110-206): Threefry white noise shaped by sqrt(P(k)),
1LPT and 2LPT displacements and the k^2 field.  Here:

* ``EH98NoWiggle``: the Eisenstein & Hu (1998, ApJ 496, 605) eqs. 26-31
  no-wiggle linear P(k), n_s tilt, sigma_8 normalised with the real-space
  top-hat (MassFunction.jl:29-42 convention), divided by (2 pi)^3 so that
  <delta^2> = int P d^3k, the field-generator convention of
  RandomField.jl:136-175; ``pk_table`` samples it into the two columns the
  reference's P(k) file would hold;
* ``make_tile``: the tile's fields generated on the GPU by the library
  (pp_make_tile_fields) into torch tensors;
* ``filter_bank``: log-spaced top-hat radii in the drivers' order.

Arrays use the reference's memory layout: a Julia ``Array{Float32,3}`` of size
(n, n, n) is x-fastest, i.e. a C-order array of shape (n, n, n) indexed
[k, j, i].
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np


class EH98NoWiggle:
    """Eisenstein & Hu 1998 no-wiggle linear P(k) in the field-generator
    convention; k in h/Mpc, P in (Mpc/h)^3 / (2 pi)^3."""

    def __init__(self, Om=0.315, Ob=0.049, h=0.674, ns=0.965, sigma8=0.808, Tcmb=2.7255):
        self.Om, self.Ob, self.h, self.ns, self.sigma8 = Om, Ob, h, ns, sigma8
        omh2, obh2 = Om * h * h, Ob * h * h
        self.theta = Tcmb / 2.7
        self.fb = Ob / Om
        self.s = 44.5 * math.log(9.83 / omh2) / math.sqrt(1.0 + 10.0 * obh2 ** 0.75)   # Mpc
        fb = self.fb
        self.alpha = 1.0 - 0.328 * math.log(431.0 * omh2) * fb + 0.38 * math.log(22.3 * omh2) * fb * fb
        self.A = 1.0
        s8 = self.sigma_R(8.0)
        self.A = (sigma8 / s8) ** 2

    def transfer(self, k):
        k = np.asarray(k, np.float64)
        ks = k * self.h * self.s                                        # k in 1/Mpc times s
        gamma = self.Om * self.h * (self.alpha + (1.0 - self.alpha) / (1.0 + (0.43 * ks) ** 4))
        q = k * self.theta ** 2 / gamma
        L0 = np.log(2.0 * math.e + 1.8 * q)
        C0 = 14.2 + 731.0 / (1.0 + 62.5 * q)
        return L0 / (L0 + C0 * q * q)

    def p_std(self, k):
        """Standard-convention P(k): sigma^2(R) = 1/(2 pi^2) int k^2 P W^2 dk."""
        k = np.asarray(k, np.float64)
        return self.A * k ** self.ns * self.transfer(k) ** 2

    def sigma_R(self, R):
        lk = np.linspace(math.log(1e-5), math.log(1e3), 20001)
        k = np.exp(lk)
        x = k * R
        w = np.where(x > 1e-4, 3.0 * (np.sin(x) - x * np.cos(x)) / np.where(x > 1e-4, x, 1.0) ** 3,
                     1.0 - x * x / 10.0)
        integrand = k ** 3 * self.p_std(k) * w * w / (2.0 * math.pi ** 2)
        return math.sqrt(np.trapezoid(integrand, lk))

    def __call__(self, k):
        return self.p_std(k) / (2.0 * math.pi) ** 3


_PK_CACHE: dict = {}


def pk_table(pk=None, kmin: float = 1e-4, kmax: float = 1e2, npts: int = 2001):
    """(k, P) columns of a P(k) file sampled from `pk` (default EH98 with the
    configs' cosmology): npts log-spaced k in [kmin, kmax] h/Mpc.  Between
    rows the reference's log-log interpolation holds P to ~1e-6."""
    key = (id(pk) if pk is not None else None, kmin, kmax, npts)
    t = _PK_CACHE.get(key)
    if t is None:
        f = pk or EH98NoWiggle()
        k = np.exp(np.linspace(math.log(kmin), math.log(kmax), npts))
        t = (k, np.asarray(f(k), np.float64))
        if pk is None:
            _PK_CACHE[key] = t
    return t


def filter_bank(nfilt: int, rmin: float, rmax: float) -> np.ndarray:
    """Log-spaced top-hat radii, returned in the driver's order: descending
    (Pipeline.jl:81-82)."""
    r = np.exp(np.linspace(math.log(rmin), math.log(rmax), nfilt))
    return np.sort(r)[::-1].copy()


class TileInputs:
    """The fields one tile of the path consumes (PeakGrid, RadialShell.jl:45-56)."""

    def __init__(self, n, boxsize, delta, eta, eta2, lapd):
        self.n, self.boxsize = n, boxsize
        self.alatt = boxsize / n
        self.delta = delta
        self.eta, self.eta2, self.lapd = eta, eta2, lapd


def make_tile(ctx, n: int, boxsize: float, seed: int = 42, ilpt: int = 2, want_lapd: bool = True,
              pk=None) -> TileInputs:
    """run_tile's fields of a periodic tile (Pipeline.jl:110-206) generated on
    ctx's GPU (pp_make_tile_fields) into CUDA float32 tensors; pk is a
    callable P(k) (default EH98), tabulated as the reference's P(k) file."""
    import torch

    from . import _lib as L
    dev = torch.device("cuda", ctx.device)
    new = lambda: torch.empty((n, n, n), dtype=torch.float32, device=dev)
    delta = new()
    eta = (new(), new(), new())
    eta2 = (new(), new(), new()) if ilpt >= 2 else None
    lapd = new() if want_lapd else None
    k, P = pk_table(pk)
    tab = L.PkTable(len(k), k.ctypes.data, P.ctypes.data)
    p = lambda t: t.data_ptr() if t is not None else None
    f = L.TileFields(p(delta), *(p(a) for a in eta), *((p(a) for a in eta2) if eta2 else (None, None, None)), p(lapd))
    ctx._call(lambda: ctx.lib.pp_make_tile_fields(ctx.h, n, float(boxsize), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                                  C.byref(tab), C.byref(f)), True)
    return TileInputs(n, boxsize, delta, eta, eta2, lapd)
