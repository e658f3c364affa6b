"""Pins for the oracle's restatement of run_tile's input fields
(oracle/fields.py: generate_grf, displacements_1lpt / _2lpt,
compute_laplacian; src/InitialConditions/RandomField.jl:123-175,
src/InitialConditions/LPT.jl:21-191, src/Pipeline.jl:35-51).  CPU only.

Plane waves below the Nyquist frequency go through each transfer in closed
form: the amplitude sqrt(P(k) dk^3 n^3) of a noise mode, psi = -(k_d / k^2)
sin for the i k_d / k^2 kernel (the sign of i k), a zero second-order source
for one plane wave and cos x cos y for two perpendicular ones, and k^2 cos
divided by n^3 once more for the reference's k^2 field.
"""
import math

import numpy as np
import pytest

from oracle import fields as F


def _pk():
    k = np.logspace(-3, 2, 400)
    return F.PkTable(k, 1e4 * k / (1.0 + (k / 0.02) ** 2) ** 1.5)


def _wave(n, m):
    i = np.arange(n)
    return (2 * math.pi / n) * (m[0] * i[None, None, :] + m[1] * i[None, :, None] + m[2] * i[:, None, None])


def test_grf_amplitude_of_a_noise_mode(monkeypatch):
    n, L, m = 16, 40.0, (2, 1, 3)
    ph = _wave(n, m)
    monkeypatch.setattr(F, "fill_noise_threefry", lambda n_, seed: np.cos(ph).astype(np.float32))
    pk = _pk()
    dk = 2 * math.pi / L
    k0 = dk * math.sqrt(sum(q * q for q in m))
    amp = math.sqrt(pk(np.array([k0]))[0] * dk ** 3 * n ** 3)
    np.testing.assert_allclose(F.generate_grf(n, L, 1, pk), amp * np.cos(ph), atol=2e-6 * amp)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_1lpt_of_a_plane_wave(d):
    n, L, m = 16, 24.0, (3, 2, 1)
    ph = _wave(n, m)
    dk = 2 * math.pi / L
    kv = [dk * q for q in m]
    k2 = sum(v * v for v in kv)
    psi = F.displacements_1lpt(F._rfft(np.cos(ph).astype(np.float32)), n, L)
    np.testing.assert_allclose(psi[d], -(kv[d] / k2) * np.sin(ph), atol=2e-6)


def test_2lpt_of_plane_waves():
    n, L = 16, 32.0
    one = np.cos(_wave(n, (2, 0, 0))).astype(np.float32)
    for p in F.displacements_2lpt(F._rfft(one), n, L):
        assert np.abs(p).max() < 1e-5
    X, Y = _wave(n, (2, 0, 0)), _wave(n, (0, 3, 0))
    two = (np.cos(X) + np.cos(Y)).astype(np.float32)
    kx, ky = 2 * math.pi * 2 / L, 2 * math.pi * 3 / L
    k2 = kx * kx + ky * ky
    psi = F.displacements_2lpt(F._rfft(two), n, L)
    np.testing.assert_allclose(psi[0], (kx / k2) * np.sin(X) * np.cos(Y), atol=2e-5)
    np.testing.assert_allclose(psi[1], (ky / k2) * np.cos(X) * np.sin(Y), atol=2e-5)
    assert np.abs(psi[2]).max() < 1e-5


def test_laplacian_field_of_a_plane_wave():
    n, L, m = 16, 24.0, (3, 0, 1)
    ph = _wave(n, m)
    k2 = (2 * math.pi / L) ** 2 * sum(q * q for q in m)
    got = F.compute_laplacian(F._rfft(np.cos(ph).astype(np.float32)), n, L)
    np.testing.assert_allclose(got, k2 * np.cos(ph) / n ** 3, atol=2e-5 * k2 / n ** 3)
