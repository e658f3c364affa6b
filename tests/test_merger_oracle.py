"""Pins for the oracle's restatement of merge_catalog
(src/Merger/Merger.jl:21-79, src/Merger/Exclusion.jl).  CPU only.

* the reference's own cases (test/test_merger.jl): sphere_overlap's special
  cases and its equal-sphere closed form (cap height r/2), exclusion of a
  halo inside a larger one, the equal-radius tie (the earlier halo of the
  stable order wins), symmetric reduction of an overlapping pair, and no
  change without overlap;
* brute force on random catalogues, without the spatial hash: a halo
  survives exclusion iff no surviving halo earlier in the descending-radius
  order has it inside its radius, and its volume deficit is the sum of its
  caps over every surviving partner at |r_i - r_j| < d < r_i + r_j.
"""
import math

import numpy as np
import pytest

from oracle import merger as OM


def test_sphere_overlap_cases():
    assert OM.sphere_overlap(10.0, 2.0, 3.0) == (0.0, 0.0)
    assert OM.sphere_overlap(5.0, 2.0, 3.0) == (0.0, 0.0)
    assert OM.sphere_overlap(0.5, 3.0, 1.0) == (0.0, 0.0)
    assert OM.sphere_overlap(0.0, 2.0, 2.0) == (0.0, 0.0)
    r = 2.0
    v1, v2 = OM.sphere_overlap(r, r, r)
    h = r / 2
    assert v1 == pytest.approx((math.pi / 3) * h * h * (3 * r - h), rel=1e-15) and v1 == v2
    v1, v2 = OM.sphere_overlap(3.0, 3.0, 1.0)
    assert 0 < v1 < v2


def test_reference_merge_cases():
    alive, r = OM.merge([0.0, 0.3, 10.0], [0.0, 0.3, 10.0], [0.0, 0.3, 10.0], [2.0, 0.5, 1.0])
    assert list(alive) == [True, False, True]
    alive, r = OM.merge([0.0, 0.5], [0.0, 0.0], [0.0, 0.0], [2.0, 2.0])
    assert list(alive) == [True, False]
    alive, r = OM.merge([0.0, 3.0], [0.0, 0.0], [0.0, 0.0], [2.0, 2.0])
    assert list(alive) == [True, True] and r[0] < 2.0 and r[0] == r[1]
    alive, r = OM.merge([0.0, 100.0], [0.0, 0.0], [0.0, 0.0], [2.0, 3.0])
    assert list(alive) == [True, True] and list(r) == [2.0, 3.0]


def _brute(x, y, z, r):
    n = len(r)
    order = sorted(range(n), key=lambda i: -r[i])
    alive = np.zeros(n, bool)
    done = []
    for i in order:
        ok = True
        for k in done:
            if alive[k] and math.dist((x[i], y[i], z[i]), (x[k], y[k], z[k])) < r[k]:
                ok = False
                break
        alive[i] = ok
        done.append(i)
    dV = np.zeros(n)
    idx = np.flatnonzero(alive)
    for i in idx:
        for j in idx:
            if i == j:
                continue
            d = math.dist((x[i], y[i], z[i]), (x[j], y[j], z[j]))
            if abs(r[i] - r[j]) < d < r[i] + r[j]:
                dV[i] += OM.sphere_overlap(d, r[i], r[j])[0]
    f = 4 * math.pi / 3
    vol = f * np.asarray(r) ** 3 - dV
    new_r = np.where(alive & (vol > 0), np.cbrt(np.maximum(vol, 0) / f), 0.0)
    return alive & (vol > 0), new_r


@pytest.mark.parametrize("seed,n", [(1, 300), (2, 400), (3, 200)])
def test_merge_against_brute_force(seed, n):
    # clustered halos in a wide box (the hash's 256 cells stay small next to
    # the radii, so the restatement's cell walks stay short)
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 200.0, (12, 3))
    pos = centres[rng.integers(0, 12, n)] + rng.normal(0, 4.0, (n, 3))
    x, y, z = (pos[:, q].astype(np.float32) for q in range(3))
    r = np.minimum(rng.pareto(2.5, n) + 1.0, 3.0).astype(np.float32) * 1.2
    r[:5] = r[5]                     # some exact ties
    alive, nr = OM.merge(x, y, z, r)
    ba, br = _brute([float(v) for v in x], [float(v) for v in y], [float(v) for v in z], [float(v) for v in r])
    assert 0 < alive.sum() < n
    np.testing.assert_array_equal(alive, ba)
    np.testing.assert_allclose(nr, br, rtol=1e-12, atol=0)
