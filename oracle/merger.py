"""CPU restatement of the catalogue merger (TEST INFRASTRUCTURE ONLY).

The oracle for the library's merger (pp_merge_halos): only tests/ may import
it.  It restates merge_catalog (src/Merger/Merger.jl:21-79) with the
reference's two passes, literally and in Python, over the reference's
256^3 linked-list spatial hash (src/Merger/Exclusion.jl:14-62):

* lagrangian_exclusion! (Exclusion.jl:98-144): halos in descending radius
  order (stable); each survivor i kills every surviving j != i with
  r_i >= r_j whose centre lies closer than r_i;
* volume_reduction! (Exclusion.jl:146-222): every surviving pair at
  |r_i - r_j| < d < r_i + r_j adds the sphere-sphere cap volumes of
  sphere_overlap (Exclusion.jl:64-96, the Wolfram formula) to both halos,
  once from each side; the deficits are halved, then
  r_new = ((4 pi / 3 r^3 - dV) / (4 pi / 3))^(1/3), and a halo whose volume
  goes to zero or below is dropped.
"""
from __future__ import annotations

import math

import numpy as np

NC = 256


class SpatialHash:
    """build_hash (Exclusion.jl:33-52): head-of-chain per cell, linked list
    per halo (1-based halo numbers, 0 = end); the last halo inserted heads
    its cell."""

    def __init__(self, x, y, z, domain_min, domain_max):
        L = max(domain_max[q] - domain_min[q] for q in range(3))
        self.cell_size = L / NC
        self.origin = domain_min
        self.hoc = {}
        self.ll = [0] * (len(x) + 1)
        for m in range(1, len(x) + 1):
            c = self.cell(x[m - 1], y[m - 1], z[m - 1])
            self.ll[m] = self.hoc.get(c, 0)
            self.hoc[c] = m

    def cell(self, px, py, pz):
        """_cell_idx (Exclusion.jl:55-61), 1-based and clamped."""
        return tuple(min(max(math.floor((float(p) - self.origin[q]) / self.cell_size) + 1, 1), NC)
                     for q, p in enumerate((px, py, pz)))

    def around(self, c, dcell):
        """Halo numbers in the cells within dcell of c, in the reference's
        traversal order (z, y, x outer to inner; each chain from its head)."""
        ci, cj, ck = c
        for diz in range(-dcell, dcell + 1):
            for diy in range(-dcell, dcell + 1):
                for dix in range(-dcell, dcell + 1):
                    jx, jy, jz = dix + ci, diy + cj, diz + ck
                    if jx < 1 or jx > NC or jy < 1 or jy > NC or jz < 1 or jz > NC:
                        continue
                    j = self.hoc.get((jx, jy, jz), 0)
                    while j > 0:
                        yield j
                        j = self.ll[j]


def sphere_overlap(d: float, r1: float, r2: float):
    """sphere_overlap (Exclusion.jl:64-96)."""
    if d >= r1 + r2:
        return 0.0, 0.0
    if d <= abs(r1 - r2):
        return 0.0, 0.0
    h1 = (r2 - r1 + d) * (r2 + r1 - d) / (2.0 * d)
    h2 = (r1 - r2 + d) * (r1 + r2 - d) / (2.0 * d)
    v1 = (math.pi / 3.0) * (h1 * h1) * (3.0 * r1 - h1)
    v2 = (math.pi / 3.0) * (h2 * h2) * (3.0 * r2 - h2)
    return v1, v2


def merge(x, y, z, r):
    """merge_catalog's passes on Float64 copies of the catalogue columns
    (x, y, z, RTHL); returns (survived bool array, new radius Float64 array
    with 0 for dropped halos)."""
    x = [float(v) for v in x]
    y = [float(v) for v in y]
    z = [float(v) for v in z]
    r = [float(v) for v in r]
    n = len(r)
    if n == 0:
        return np.zeros(0, bool), np.zeros(0)
    order = sorted(range(n), key=lambda i: -r[i])    # stable, descending
    rmax = max(r)
    pad = 2.0 * rmax
    lo = (min(x) - pad, min(y) - pad, min(z) - pad)
    hi = (max(x) + pad, max(y) + pad, max(z) + pad)
    sh = SpatialHash(x, y, z, lo, hi)
    alive = [True] * n
    # pass 1: Lagrangian exclusion
    for i in order:
        if not alive[i]:
            continue
        ri, xi, yi, zi = r[i], x[i], y[i], z[i]
        dcell = math.ceil(ri / sh.cell_size)
        for j1 in sh.around(sh.cell(xi, yi, zi), dcell):
            j = j1 - 1
            if j != i and alive[j] and ri >= r[j]:
                dx, dy, dz = x[j] - xi, y[j] - yi, z[j] - zi
                if math.sqrt(dx * dx + dy * dy + dz * dz) < ri:
                    alive[j] = False
    # pass 2: volume reduction
    dV = [0.0] * n
    for i in order:
        if not alive[i]:
            continue
        ri, xi, yi, zi = r[i], x[i], y[i], z[i]
        dcell = math.ceil((ri + r[order[0]]) / sh.cell_size)
        for j1 in sh.around(sh.cell(xi, yi, zi), dcell):
            j = j1 - 1
            if j != i and alive[j]:
                rj = r[j]
                dx, dy, dz = x[j] - xi, y[j] - yi, z[j] - zi
                dist = math.sqrt(dx * dx + dy * dy + dz * dz)
                if dist < ri + rj and dist > abs(ri - rj):
                    v1, v2 = sphere_overlap(dist, ri, rj)
                    dV[i] += v1
                    dV[j] += v2
    new_r = list(r)
    f = 4.0 * math.pi / 3.0
    for i in range(n):
        if not alive[i]:
            continue
        vol = f * (r[i] * r[i] * r[i]) - 0.5 * dV[i]
        if vol <= 0.0:
            alive[i] = False
            new_r[i] = 0.0
        else:
            new_r[i] = (vol / f) ** (1.0 / 3.0)
    new_r = [v if a else 0.0 for v, a in zip(new_r, alive)]
    return np.array(alive), np.array(new_r)
