"""Analytic test cases for the outputs of analyse_peak that no reference-held
vector fixes (RadialShell.jl:706-885): R_TH interpolation, Fbarx, the strain
and its e_v / p_v, gradpk and the permuted-read gradpkf (quirk Q6), Sbar2,
the exclusion mask, d2F (quirk Q7) and zvir_half.

Each case is a field whose answers follow from the mathematics rather than
from the reference's code:

* delta is a spherical Gaussian bump about a centre on the grid diagonal plus
  a linear ramp g . (x - c).  Every full lattice shell is symmetric under the
  cube group, so the ramp adds nothing to a shell mean (Fbar is the bump's)
  and the kernel-weighted gradient of the ramp is the ramp's slope g (per
  lattice unit, alatt = 1).  The permuted read delta[v3, v1, v2] sees the
  ramp with its slopes rotated, (gx, gy, gz) -> (gy, gz, gx).
* eta_L = -s_L (x_L - c) is a uniform strain, so every kernel-weighted strain
  is diag(s) times a positive scalar: e_v = (s3 - s1) / (2 sum s),
  p_v = (s3 + s1 - 2 s2) / (2 sum s), and the reported strain is
  diag(s) Fbarx / sum s.
* The collapse table is constant (every query collapses: R_TH is the radius
  of the last shell with Fbar >= fcrit and zvir_half = value - 1) or linear
  in log10 F and constant in e and p/e (R_TH interpolates zvir linearly in r^3
  between the two shells around the crossing of 1 + z).
* eta2_x = r^2, eta2_y = 1 + r^2 / 2, eta2_z = 1/4 (Sbar2 is the mean over the
  cells inside the collapse shell), lapd = r^2 + 2 (d2F is a mean over cells
  2.. in shell order inside R_TH; the centre is excluded and the loop stops at
  the first cell outside the grid).

The expected values are computed here from those statements with numpy, not
by the oracle; the oracle and both GPU forms are compared against them.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SLOPE = (0.010, 0.020, 0.035)       # ramp of delta per lattice unit (x, y, z)
STRAIN = (0.5, 1.0, 2.0)            # eta_L = -s_L (x_L - c)


@dataclass
class Case:
    n: int
    centre: tuple          # 0-based (i, j, k)
    delta: np.ndarray      # (n, n, n) float32, [k, j, i]
    eta: tuple
    eta2: tuple
    lapd: np.ndarray
    rmax: int
    nbuff: int
    ir2min: int = 2

    @property
    def ipp(self) -> int:
        i, j, k = self.centre
        return (i + 1) + j * self.n + k * self.n * self.n


def bump_case(n=48, centre=(23, 23, 23), sigma=4.0, amp=12.0, rmax=20, nbuff=3) -> Case:
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    d = (i - centre[0], j - centre[1], k - centre[2])
    r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
    delta = amp * np.exp(-r2 / (2.0 * sigma * sigma)) + sum(g * x for g, x in zip(SLOPE, d))
    eta = tuple((-s * x).astype(np.float32) for s, x in zip(STRAIN, d))
    eta2 = (r2.astype(np.float32), (1.0 + 0.5 * r2).astype(np.float32), np.full(r2.shape, 0.25, np.float32))
    lapd = (r2 + 2.0).astype(np.float32)
    return Case(n, centre, delta.astype(np.float32), eta, eta2, lapd, rmax, nbuff)


def shells(rmax: int):
    """Distinct r^2 of the lattice ball of radius rmax, ascending."""
    r = np.arange(-rmax, rmax + 1)
    q = (r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2).ravel()
    return np.unique(q[q <= rmax * rmax])


def fbar_profile(case: Case):
    """(rad, Fbar) of shells 1..m (1-based arrays, index 0 unused), m = the
    finalised shells (all but the last, quirk Q4): Fbar is the r^3-weighted
    trapezoid of the shell means, Fbar(1) the centre value."""
    r2v = shells(case.rmax)
    ci, cj, ck = case.centre
    k, j, i = np.meshgrid(np.arange(case.n), np.arange(case.n), np.arange(case.n), indexing="ij")
    q = (i - ci) ** 2 + (j - cj) ** 2 + (k - ck) ** 2
    d = case.delta.astype(np.float64)
    m = len(r2v) - 1
    rad = np.zeros(m + 1)
    fb = np.zeros(m + 1)
    fb[1] = d[ck, cj, ci]
    prev_mean = fb[1]
    for s in range(2, m + 1):
        sel = q == r2v[s - 1]
        mean = d[sel].mean()
        ra, rb = math.sqrt(r2v[s - 2]), math.sqrt(r2v[s - 1])
        rad[s] = rb
        fb[s] = (ra ** 3 * fb[s - 1] + 0.5 * (prev_mean + mean) * (rb ** 3 - ra ** 3)) / rb ** 3
        prev_mean = mean
    return rad, fb


def crossing(fb, fcrit) -> int:
    """First shell whose mean overdensity Fbar drops below fcrit."""
    return int(np.argmax(fb[1:] < fcrit)) + 1


def linear_table(a: float, b: float, X1=-3.0, X2=3.0):
    """A 2x2x2 table whose value a + b log10 F is exact under trilinear
    interpolation for X1 <= log10 F <= X2, independent of e and p/e."""
    t = np.empty((2, 2, 2), np.float32)
    t[0] = a + b * X1
    t[1] = a + b * X2
    return t, (X1, X2, -10.0, 10.0, -1e3, 1e3)


def constant_table(c: float):
    return np.full((2, 2, 2), c, np.float32), (-3.0, 3.0, -10.0, 10.0, -1e3, 1e3)


def expect_linear(case: Case, a: float, b: float, fcrit: float, ZZon: float):
    """R_TH, Fbarx, the collapse shell and zvir_half for linear_table(a, b)."""
    rad, fb = fbar_profile(case)
    m = len(fb) - 1
    m0 = crossing(fb, fcrit)
    # the table holds Float32 end values: the line through them
    t, (X1, X2, *_) = linear_table(a, b)
    lo, hi = float(t[0, 0, 0]), float(t[1, 0, 0])
    zv = lambda s: (lo + (math.log10(fb[s]) - X1) / (X2 - X1) * (hi - lo)
                    if fb[s] > 0 and X1 <= math.log10(fb[s]) <= X2 else -1.0)
    # the collapse shell: the outermost shell inside the crossing whose zvir
    # reaches 1 + z; the shell outside it (or the crossing, forced to -1)
    # brackets the threshold
    mp = next(s for s in range(m0 - 1, 0, -1) if zv(s) >= ZZon)
    z_in = zv(mp)
    z_out = zv(mp + 1) if mp + 1 < m0 else -1.0
    if z_out > 0:
        r3 = rad[mp] ** 3 + (rad[mp + 1] ** 3 - rad[mp] ** 3) * (z_in - ZZon) / (z_in - z_out)
        R = r3 ** (1.0 / 3.0)
        t = (r3 - rad[mp] ** 3) / (rad[mp + 1] ** 3 - rad[mp] ** 3)
        Fx = fb[mp] + t * (fb[mp + 1] - fb[mp])
    else:
        R, Fx = rad[mp], fb[mp]
    zh = -1.0
    if R >= 3.0:
        vals = []
        for jj in range(10):
            rc = R * ((jj / 9.0) * 0.5 + 0.5) ** (1.0 / 3.0)
            # the first shell (of 1..m-1) beyond rcur, else the last of them
            s = next((s for s in range(1, m) if rad[s] > rc), m - 1)
            vals.append(zv(s) - 1.0)
        zh = max(vals)
    return {"RTHL": R, "Fbarx": Fx, "inside": mp, "zvir_half": zh}


def inside_cells(case: Case, R: float, frame: int = 0, exclude_centre=False):
    """Grid cells within R of the centre, inside [frame, n - frame)^3."""
    n = case.n
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ci, cj, ck = case.centre
    q = (i - ci) ** 2 + (j - cj) ** 2 + (k - ck) ** 2
    sel = np.sqrt(q.astype(np.float64)) <= R
    if frame:
        lo, hi = frame, n - frame
        sel &= (i >= lo) & (i < hi) & (j >= lo) & (j < hi) & (k >= lo) & (k < hi)
    if exclude_centre:
        sel &= q > 0
    return sel, q


def expect_common(case: Case, R: float, Fbarx: float, inside_r2: float):
    """Strain, e_v, p_v, gradients, Sbar2 and d2F of a case whose sphere of
    radius R lies inside the grid; inside_r2 = r^2 of the collapse shell."""
    s = np.array(STRAIN)
    sel_s, q = inside_cells(case, math.sqrt(inside_r2) + 1e-9)
    e2x = case.eta2[0].astype(np.float64)[sel_s].mean()
    e2y = case.eta2[1].astype(np.float64)[sel_s].mean()
    sel_d, _ = inside_cells(case, R, exclude_centre=True)
    d2F = case.lapd.astype(np.float64)[sel_d].mean()
    return {"strain": np.diag(s) * Fbarx / s.sum(), "e_v": 0.5 * (s[2] - s[0]) / s.sum(),
            "p_v": 0.5 * (s[2] + s[0] - 2 * s[1]) / s.sum(), "gradpk": np.array(SLOPE),
            "gradpkf": np.array(SLOPE)[[1, 2, 0]], "Sbar2": np.array([e2x, e2y, 0.25]), "d2F": d2F}


def d2F_walk(case: Case, R: float, di, dj, dk, r2):
    """d2F by the rule of quirk Q7 over a shell table given in its order:
    cells 2.. with r <= R, stopping at the first cell outside the grid."""
    ci, cj, ck = case.centre
    n = case.n
    tot, cnt = 0.0, 0
    for p in range(1, len(r2)):
        if math.sqrt(r2[p]) > R:
            continue
        x, y, z = ci + di[p], cj + dj[p], ck + dk[p]
        if not (0 <= x < n and 0 <= y < n and 0 <= z < n):
            break
        tot += float(case.lapd[z, y, x])
        cnt += 1
    return tot / cnt if cnt else 0.0
