"""Catalogue format and record-packing formulas (CPU only)."""
import math
import os

import numpy as np
import pytest

from oracle import catalog as OC
from peakpatch_hip import catalog


def test_pksc_roundtrip_and_layout(tmp_path):
    """write_pksc / read_pksc (Catalog.jl:53-160; test_catalog.jl:1-60):
    12-byte header + N x 11 x 4 (basic) or N x 33 x 4 (extended)."""
    basic = np.array([[1, 2, 3, .1, .2, .3, 5, -.1, -.2, -.3, 1.686],
                      [10, 20, 30, 1, 2, 3, 8, -1, -2, -3, 2.5]], np.float32)
    p = os.path.join(tmp_path, "b.pksc")
    catalog.write_pksc(p, basic, 8.0, 0.5)
    assert os.path.getsize(p) == 12 + 2 * 11 * 4
    raw = open(p, "rb").read()
    assert np.frombuffer(raw[:4], "<i4")[0] == 2
    assert np.array_equal(np.frombuffer(raw[12:], "<f4").reshape(2, 11), basic)
    rec, rmax, z = catalog.read_pksc(p)
    assert np.array_equal(rec, basic) and rmax == np.float32(8.0) and z == np.float32(0.5)
    ext = np.arange(3 * 33, dtype=np.float32).reshape(3, 33)
    catalog.write_pksc(p, ext, 98.0, 0.0)
    assert os.path.getsize(p) == 12 + 3 * 33 * 4
    rec, _, _ = catalog.read_pksc(p)
    assert rec.shape == (3, 33) and np.array_equal(rec, ext)
    catalog.write_pksc(p, np.zeros((0, 11), np.float32), 0.0, 0.0)
    rec, rmax, _ = catalog.read_pksc(p)
    assert rec.shape == (0, 11) and rmax == 0


def test_packing_velocity_formulas():
    """The velocity case of test_pipeline.jl:104-146: Sbar = (1, 2, 3),
    Sbar2 = (0.5, -0.3, 0.1), RTHL = 5, alatt = 3 at z = 0 (D = 1, so the
    1LPT velocity is Sbar and the 2LPT one is -Sbar2 (-3/7) Om_a^(-1/143))."""
    Om, OL = 0.315, 0.685
    res = {"RTHL": np.array([5.0, -1.0]), "Sbar": np.array([[1.0, 2.0, 3.0], [0, 0, 0]]),
           "Sbar2": np.array([[0.5, -0.3, 0.1], [0, 0, 0]]), "Fbarx": np.array([2.0, 0.0])}
    Om_a = Om / (Om + OL)
    v2 = -(-3.0 / 7.0 * Om_a ** (-1.0 / 143) * 1.0 ** 2)
    rec, rmax = OC.pack(res, [1.0, 2.0], [3.0, 4.0], [5.0, 6.0], [2.0, 2.0], [1.0, 1.0], [1.0, 1.0],
                        [v2, v2], [1.0, 1.0], None, 3.0, False)
    assert rec.shape == (1, 11)                         # the RTHL <= 0 peak is skipped
    assert rec[0, 6] == np.float32(15.0) and rmax == np.float32(15.0)
    np.testing.assert_array_equal(rec[0, 3:6], np.float32([1, 2, 3]))
    coeff = -3.0 / 7.0 * Om_a ** (-1.0 / 143)
    assert rec[0, 7] == pytest.approx(-0.5 * coeff, abs=1e-6)
    assert rec[0, 10] == np.float32(2.0)


def test_halo_mass_closed_form():
    """M200m = 4/3 pi (2.775e11 Om h^2) R^3 (HDF5Ext.jl:58)."""
    m = catalog.halo_mass([1.0, 2.0], 0.315, 0.674)
    base = 4.0 / 3.0 * math.pi * 2.775e11 * 0.315 * 0.674 ** 2
    np.testing.assert_allclose(m, [base, 8 * base], rtol=1e-15)
