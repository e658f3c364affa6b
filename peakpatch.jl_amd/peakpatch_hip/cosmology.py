"""Host-side Float64 cosmology the driver needs around the hot path: the
linear growth factor for the per-peak collapse threshold fcrit(z) =
1.686 D(z), and the light-cone redshift of a peak from its comoving distance.

Restates src/Cosmology/Cosmology.jl:68-372 (Dlinear_tables RK4 integration,
4-point Lagrange lookup, chi(z), chi->z table + bisection, peak_redshift) and
fsc_of_z (src/HaloFinder/RadialShell.jl:191-196), vectorised over peaks.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class CosmologyParams:
    Om: float = 0.315   # total matter
    OB: float = 0.049
    OL: float = 0.685
    h: float = 0.674
    ns: float = 0.965
    s8: float = 0.808


def _deriv(y, a, h, Omnr, Omvac, Omcurv, OmX, OmB):
    # Cosmology.jl:123-146 with no radiation, no hot dark matter, no decay
    a3 = a * a * a
    hub = Omnr + Omvac * a3 + Omcurv * a
    d0 = 1.0 / (h * math.sqrt(hub / a3))
    fb = y[2]
    q = hub
    qs = 0.5 / q
    q_val = qs * (Omnr - 2.0 * Omvac * a3)
    fnr = (OmX + OmB) / Omnr
    if OmB != 0.0 and a <= 1.0e-3:
        fnr = OmX / Omnr
    qf = qs * (Omnr * fnr)
    return (d0, fb, 3.0 * qf - fb * fb - (1.0 - q_val) * fb, 0.0, d0 / a)


class GrowthTables:
    """Dlinear_tables / Dlinear_ab (Cosmology.jl:148-293)."""

    def __init__(self, c: CosmologyParams):
        Omnr, Omvac, Omcurv, OmX = c.Om, c.OL, 1.0 - c.Om - c.OL, c.Om - c.OB
        amaxtab, amin_target, ainit, nmax = 2.0, 1.0 / 1001.0, 1.0e-7, 2000
        a, alogn = ainit, math.log(ainit)
        hh = (math.log(amaxtab) - alogn) / nmax
        t0 = 1.0 / c.h / math.sqrt(Omnr)
        ya = math.sqrt(a)
        y = [t0 * (2.0 / 3.0 * ya ** 3), math.log(a),
             0.25 * (math.sqrt(24.0 * OmX / Omnr + 1.0) - 1.0), 0.0, t0 * 2.0 * ya]
        args = (c.h, Omnr, Omvac, Omcurv, OmX, c.OB)
        at, Dt, ft = [], [], []
        lnorm, bnorm, ip0 = True, 1.0, 0
        for _ in range(nmax):
            w1 = _deriv(y, a, *args)
            ah = math.exp(alogn + hh * 0.5)
            w2 = _deriv([y[q] + hh * w1[q] * 0.5 for q in range(5)], ah, *args)
            w3 = _deriv([y[q] + hh * w2[q] * 0.5 for q in range(5)], ah, *args)
            alogn += hh
            a = math.exp(alogn)
            w4 = _deriv([y[q] + hh * w3[q] for q in range(5)], a, *args)
            y = [y[q] + hh / 6.0 * (w1[q] + w4[q] + 2.0 * (w2[q] + w3[q])) for q in range(5)]
            if lnorm and a >= 0.001:
                lnorm, bnorm = False, math.exp(y[1]) / a
            if amin_target <= a <= amaxtab:
                at.append(a); Dt.append(math.exp(y[1]) / a); ft.append(y[2])
                if a >= 1.0 and ip0 == 0:
                    ip0 = len(at)
        if ip0 > 1:
            ai = (at[ip0 - 1] - 1.0) / (at[ip0 - 1] - at[ip0 - 2])
            Dg = (Dt[ip0 - 1] * (1.0 - ai) + Dt[ip0 - 2] * ai) / bnorm
            fac = 1.0 / (Dg * bnorm)
        else:
            fac = 1.0 / bnorm
        self.atab = np.array(at)
        self.Datab = np.array(Dt) * fac
        self.dlnD = np.array(ft)

    def D(self, a):
        """Dlinear(a) = a * D_a (first output of Dlinear_ab), vectorised."""
        a = np.atleast_1d(np.asarray(a, np.float64))
        return self.D_a(a) * a

    def D_a(self, a):
        """D_a, the interpolated table value (third output of Dlinear_ab,
        Cosmology.jl:273-293), vectorised; the drivers' velocity factor."""
        a = np.atleast_1d(np.asarray(a, np.float64))
        at, Dt = self.atab, self.Datab
        n = len(at)
        kk = np.searchsorted(at, a, side="left")          # _hunt: at[kk-1] < a <= at[kk] (1-based kk)
        out = np.empty_like(a)
        inside = (a > at[0]) & (a < at[-1])
        k = kk[inside]                                      # 1-based index of the last at < a
        x = a[inside]
        two = (k < 3) | (k >= n)                            # k == 2 would index atab[0] in the reference
        lo = np.where(two, np.maximum(k - 1, 1) - 1, k - 3)
        lo = np.clip(lo, 0, n - 4)
        off2 = np.maximum(k - 1, 1) - 1 - lo                # 2-point nodes relative to lo
        Da = np.empty_like(x)
        # 4-point Lagrange (_lagrint_2col, Cosmology.jl:85-101)
        xa = np.stack([at[lo + q] for q in range(4)])
        va = np.stack([Dt[lo + q] for q in range(4)])
        D4 = np.zeros_like(x)
        for p in range(4):
            w = np.ones_like(x)
            for q in range(4):
                if q != p:
                    w = w * (x - xa[q]) / (xa[p] - xa[q])
            D4 = D4 + w * va[p]
        # 2-point rule
        i0 = lo + off2
        x0, x1 = at[i0], at[np.minimum(i0 + 1, n - 1)]
        a1 = (x - x1) / (x0 - x1)
        D2 = a1 * Dt[i0] + (1.0 - a1) * Dt[np.minimum(i0 + 1, n - 1)]
        out[inside] = np.where(two, D2, D4)
        out[a >= at[-1]] = Dt[-1]
        out[a <= at[0]] = Dt[0]
        return out

    def fsc_of_z(self, z):
        """fsc_of_z (RadialShell.jl:191-196): 1.686 * D(a = 1/(1+z))."""
        z = np.asarray(z, np.float64)
        return 1.686 * self.D(1.0 / (1.0 + z)).reshape(z.shape)


_GL64 = np.polynomial.legendre.leggauss(64)


def chi(z: float, c: CosmologyParams) -> float:
    """chi(z) (Cosmology.jl:21-24): (2.998e5/100) h int_0^z dz/E(z) [the
    reference's h factor kept]; 64-point Gauss-Legendre per unit redshift."""
    if z <= 0.0:
        return 0.0
    nseg = max(1, int(math.ceil(z)))
    x, w = _GL64
    tot = 0.0
    for s in range(nseg):
        a, b = z * s / nseg, z * (s + 1) / nseg
        zz = 0.5 * (b - a) * x + 0.5 * (a + b)
        tot += 0.5 * (b - a) * np.sum(w / np.sqrt(c.Om * (1 + zz) ** 3 + c.OL))
    return (2.998e5 / 100.0) * c.h * tot


class ChiToZ:
    """build_chi_to_z / chi_to_z (Cosmology.jl:333-358)."""

    def __init__(self, c: CosmologyParams, z_max: float = 10.0, npts: int = 2000):
        self.zs = z_max * np.arange(npts) / (npts - 1)
        self.chis = np.array([chi(z, c) for z in self.zs])
        self.z_max = z_max

    def __call__(self, chi_val):
        chi_val = np.asarray(chi_val, np.float64)
        zs, cs = self.zs, self.chis
        hi = np.searchsorted(cs, chi_val, side="right")     # chis[lo] <= chi < chis[hi]
        hi = np.clip(hi, 1, len(cs) - 1)
        lo = hi - 1
        frac = (chi_val - cs[lo]) / (cs[hi] - cs[lo])
        z = zs[lo] + frac * (zs[hi] - zs[lo])
        z = np.where(chi_val <= cs[0], 0.0, z)
        return np.where(chi_val >= cs[-1], self.z_max, z)

    def peak_redshift(self, obs, x, y, z):
        """peak_redshift (Cosmology.jl:366-372)."""
        d = np.sqrt((np.asarray(x) - obs[0]) ** 2 + (np.asarray(y) - obs[1]) ** 2 + (np.asarray(z) - obs[2]) ** 2)
        return self(d)


def ir2min(Rf, alatt: float):
    """Per-peak ir2min of the drivers (Pipeline.jl:222-223)."""
    Rf = np.asarray(Rf, np.float64)
    a = np.floor((1.75 * Rf / alatt) ** 2)
    b = math.floor((40.0 / alatt - 1.0) ** 2)
    return np.minimum(a, b).astype(np.int64)
