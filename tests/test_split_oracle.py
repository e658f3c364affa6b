"""Pins for the oracle's restatement of the multi-resolution split
(run_multitile_split, src/MultiResolution.jl:41-314, 523-850).  CPU only.

Each check is fixed by something other than the oracle's own formula:
* downsample_noise: the reference's literal sequential loop on a tiny grid;
* the Hoffman-Ribak residual: zero sum over every whole coarse block;
* Catmull-Rom: weights sum to one, node values reproduced exactly at integer
  coarse coordinates, linear fields reproduced away from the periodic wrap;
* isolated_convolve: explicit DFT sums (no FFT library) on a tiny padded box,
  with the c2r rule of the inverse (weight 2 for interior x frequencies);
* periodic_convolve: a plane wave comes back as amp(k0) cos (kernel 0) and
  -amp(k0) k_d / k0^2 sin (kernel 1), which fixes the sign of i k;
* split_2lpt: a single plane wave has zero second-order source; two
  perpendicular waves give phi_xx phi_yy = cos cos;
* split_laplacian: a plane wave comes back times k^2.
"""
import math

import numpy as np
import pytest

from oracle import fields as F


def _pk():
    k = np.logspace(-3, 2, 400)
    return F.PkTable(k, 1e4 * k / (1.0 + (k / 0.02) ** 2) ** 1.5)


def test_downsample_matches_the_sequential_loop():
    N, M, seed = 6, 2, 11
    b = N // M
    ref = np.zeros((M, M, M))
    for K in range(M):
        for J in range(M):
            for I in range(M):
                val = 0.0
                for dk in range(b):
                    for dj in range(b):
                        for di in range(b):
                            gi, gj, gk = I * b + di + 1, J * b + dj + 1, K * b + dk + 1
                            idx = gi + (gj - 1) * N + (gk - 1) * N * N
                            g1, g2 = F.threefry_gaussian(seed, np.uint64((idx - 1) >> 1))
                            val += float(g1 if (idx - 1) % 2 == 0 else g2)
                ref[K, J, I] = val / math.sqrt(float(b) ** 3)
    got = F.downsample_noise(N, M, seed)
    np.testing.assert_array_equal(got, ref.astype(np.float32))


def test_residual_has_zero_block_means():
    # one tile covering the whole grid (nbuff 0): every coarse block lies inside it
    g = F.SplitGeom(12, 0, 1, 3, 1.0)
    cn = F.downsample_noise(g.N, g.M, 5)
    r = F.extended_residual(g, 1, 1, 1, 5, cn).astype(np.float64)
    b = g.block
    sums = r.reshape(g.M, b, g.M, b, g.M, b).sum(axis=(1, 3, 5))
    assert np.abs(sums).max() < b ** 3 * 2e-6
    # and the fine noise itself does not (the residual is not trivially zero)
    assert np.abs(r).max() > 0.5


def test_catmull_rom_weights():
    t = np.linspace(0, 1, 33, dtype=np.float32)[:-1]
    w = F.catmull_rom(t)
    np.testing.assert_allclose(sum(x.astype(np.float64) for x in w), 1.0, atol=1e-6)
    w0 = F.catmull_rom(np.zeros(1, np.float32))
    assert [float(x[0]) for x in w0] == [0.0, 1.0, 0.0, 0.0]


def test_interpolation_nodes_and_linear_fields():
    # block 5 (odd): cells gi = 5 m - 2 sit exactly on coarse node m
    g = F.SplitGeom(20, 0, 1, 4, 1.0)
    rng = np.random.default_rng(0)
    c = rng.standard_normal((4, 4, 4)).astype(np.float32)
    out = F.interpolate_to_tile(c, g, 1, 1, 1)
    nodes = np.arange(4) * 5 + 2                      # 0-based cells on the nodes
    np.testing.assert_array_equal(out[np.ix_(nodes, nodes, nodes)], c)
    # a field linear in the coarse x index, read where no tap wraps
    g = F.SplitGeom(48, 0, 1, 8, 1.0)
    lin = np.broadcast_to((np.arange(8, dtype=np.float32) * 0.25 + 1.0)[None, None, :], (8, 8, 8)).copy()
    out = F.interpolate_to_tile(lin, g, 1, 1, 1)
    gi = np.arange(1, 49)
    fx = (gi - 0.5) / 6 + 0.5                          # 1-based coarse coordinate
    ok = (np.floor(fx) >= 2) & (np.floor(fx) <= 6)     # taps floor-1 .. floor+2 inside 1..8
    want = (fx - 1.0) * 0.25 + 1.0
    np.testing.assert_allclose(out[3, 3, ok], want[ok], rtol=0, atol=2e-6)


def _dft_isolated(noise, pk, boxsize_local, n, kernel, dim):
    """Explicit DFT sums: forward over the padded box, half-spectrum factor,
    then the c2r inverse (x frequencies 0 and n2/2 weight 1, others 2)."""
    n_in = noise.shape[0]
    n2 = 2 * n_in
    pad = np.zeros((n2, n2, n2))
    pad[:n_in, :n_in, :n_in] = noise
    x = np.arange(n2)
    dk = 2.0 * math.pi / (n2 * (boxsize_local / n))
    fx = np.arange(n2 // 2 + 1)
    fs = np.where(x < (n2 + 1) // 2, x, x - n2)
    ex = np.exp(-2j * math.pi * np.outer(fx, x) / n2)        # [kx, x]
    ey = np.exp(-2j * math.pi * np.outer(x, x) / n2)         # [ky, y]
    H = np.einsum("zyx,ax,by,cz->cba", pad, ex, ey, ey)      # [kz, ky, kx]
    kx = (fx * dk)[None, None, :]
    ky = (fs * dk)[None, :, None]
    kz = (fs * dk)[:, None, None]
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.sqrt(pk(np.sqrt(k2)) * dk ** 3 * n2 ** 3)
        fac = amp if kernel == 0 else 1j * amp * (kx, ky, kz)[dim] / k2
    fac = np.where(k2 == 0, 0, fac)
    G = H * fac
    w = np.full(n2 // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    out = np.einsum("cba,a,ax,by,cz->zyx", G, w, ex.conj(), ey.conj(), ey.conj()).real / n2 ** 3
    return out[:n, :n, :n]


@pytest.mark.parametrize("kernel,dim", [(0, 0), (1, 0), (1, 1), (1, 2)])
def test_isolated_convolve_against_explicit_dft(kernel, dim):
    rng = np.random.default_rng(3)
    noise = rng.standard_normal((3, 3, 3)).astype(np.float32)
    pk = _pk()
    got = F.isolated_convolve(noise, pk, 3.0 * 7.5, 3, kernel, dim)
    ref = _dft_isolated(noise.astype(np.float64), pk, 3.0 * 7.5, 3, kernel, dim)
    scale = np.abs(ref).max()
    assert scale > 0
    np.testing.assert_allclose(got, ref, rtol=0, atol=3e-6 * scale)


@pytest.mark.parametrize("kernel,dim", [(0, 0), (1, 0), (1, 1), (1, 2)])
def test_periodic_convolve_plane_wave(kernel, dim):
    n, L = 16, 40.0
    m = (2, 1, 3)                                    # mode numbers along x, y, z
    i = np.arange(n)
    kvec = [2 * math.pi * q / L for q in m]
    ph = (2 * math.pi / n) * (m[0] * i[None, None, :] + m[1] * i[None, :, None] + m[2] * i[:, None, None])
    noise = np.cos(ph).astype(np.float32)
    pk = _pk()
    k0 = math.sqrt(sum(q * q for q in kvec))
    dk = 2 * math.pi / L
    amp = math.sqrt(pk(np.array([k0]))[0] * dk ** 3 * n ** 3)
    got = F.periodic_convolve(noise, L, pk, kernel, dim)
    want = amp * np.cos(ph) if kernel == 0 else -amp * kvec[dim] / k0 ** 2 * np.sin(ph)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-6 * np.abs(want).max())


def test_split_2lpt_plane_waves():
    n, L = 16, 32.0
    i = np.arange(n)
    a = np.cos(2 * math.pi * 2 * i / n)
    one = np.broadcast_to(a[None, None, :], (n, n, n)).astype(np.float32)
    for psi in F.split_2lpt(one, n, L):
        assert np.abs(psi).max() < 1e-5
    b = np.cos(2 * math.pi * 3 * i / n)
    two = (a[None, None, :] + b[None, :, None] + 0 * i[:, None, None]).astype(np.float32)
    # src = phi_xx phi_yy = cos_x cos_y, so psi2_x = -i kx/k^2 of it: for the
    # (2, 3, 0) and (2, -3, 0) modes, k^2 = kx^2 + ky^2 on both, and -i e^(i t)
    # has real part sin t
    kx, ky = 2 * math.pi * 2 / L, 2 * math.pi * 3 / L
    X, Y = np.meshgrid(2 * math.pi * 2 * i / n, 2 * math.pi * 3 * i / n, indexing="xy")
    psi = F.split_2lpt(two, n, L)
    k2 = kx * kx + ky * ky
    want_x = (kx / k2) * np.sin(X) * np.cos(Y)
    want_y = (ky / k2) * np.cos(X) * np.sin(Y)
    np.testing.assert_allclose(psi[0][0], want_x, atol=2e-5)
    np.testing.assert_allclose(psi[1][0], want_y, atol=2e-5)
    assert np.abs(psi[2]).max() < 1e-5


def test_split_laplacian_plane_wave():
    n, L = 16, 24.0
    i = np.arange(n)
    ph = 2 * math.pi * (3 * i[None, None, :] + 0 * i[None, :, None] + 1 * i[:, None, None]) / n
    d = np.cos(ph).astype(np.float32)
    k2 = (2 * math.pi / L) ** 2 * (9 + 1)
    np.testing.assert_allclose(F.split_laplacian(d, n, L), k2 * np.cos(ph), atol=2e-5 * k2)


def test_split_tile_end_to_end_is_finite():
    g = F.SplitGeom(16, 2, 2, 4, 2.0)
    t = F.split_tile(g, 1, 2, 1, 9, _pk())
    for a in (t.delta, *t.eta, *t.eta2, t.lapd):
        assert a.shape == (16, 16, 16) and np.isfinite(a).all() and np.abs(a).max() > 0
