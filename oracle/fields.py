"""CPU restatement of the tile input fields (TEST INFRASTRUCTURE ONLY).

The oracle for the library's tile synthesis (pp_make_tile_fields): only
tests/, __graft_entry__.smoke() and bench.py's CPU baseline may import it.

It restates, literally and in numpy:
* Threefry-2x64-20 (Salmon et al. 2011, Random123) and fill_noise_threefry!
  (src/InitialConditions/RandomField.jl:46-81): cell idx (1-based, x fastest)
  takes the first normal of counter pair (idx-1) >> 1 when idx-1 is even,
  else the second; uniforms from the top 53 bits, Box-Muller in Float64;
* load_pk (src/InitialConditions/PowerSpectrum.jl:11-23): P(k) by linear
  interpolation of log P against log k, extrapolated linearly (Line());
* generate_grf / _convolve_noise! (RandomField.jl:123-175): rfft of the noise,
  every mode times sqrt(P(k) dk^3 n^3) in Float64 (k = 0 zeroed), the product
  rounded to the Float32 array, normalised irfft;
* displacements_1lpt / displacements_2lpt (src/InitialConditions/LPT.jl:
  21-191): psi_i = i k_i / k^2 delta_k with k = 0 and the Nyquist planes
  zeroed; phi_ij = -k_i k_j / k^2 delta_k, the Float32 source
  phi11 phi22 - phi12^2 + phi11 phi33 - phi13^2 + phi22 phi33 - phi23^2
  evaluated left to right, psi2_i = -i k_i / k^2 src_k;
* _compute_laplacian (src/Pipeline.jl:35-51): k^2 delta_k with k in
  Float32, normalised irfft divided by n^3 once more.

and run_multitile_split's multi-resolution split of the same field
(src/MultiResolution.jl:41-314, 523-850): the coarse block-sum noise, the
periodic coarse convolutions, the Hoffman-Ribak residual, the isolated
zero-padded convolution, Catmull-Rom interpolation and the tile-local 2LPT
and k^2 fields (the section at the end of this file).

Arrays use the reference's memory layout: a Julia Array{Float32,3} of size
(n, n, n) is x-fastest, i.e. a C-order numpy array of shape (n, n, n) indexed
[k, j, i].  Transforms are scipy's pocketfft in Float64 on Float32 data.
"""
from __future__ import annotations

import math

import numpy as np

_ROT = (16, 42, 12, 31, 16, 32, 24, 21)
_PARITY = np.uint64(0x1BD11BDAA9FC1A22)


def _rotl(x, r: int):
    return (x << np.uint64(r)) | (x >> np.uint64(64 - r))


def threefry2x64_20(ctr0, ctr1, key0, key1):
    """Threefry-2x64-20 block function; uint64 scalars or arrays."""
    with np.errstate(over="ignore"):
        k0 = np.asarray(key0, np.uint64)
        k1 = np.asarray(key1, np.uint64)
        ks = (k0, k1, _PARITY ^ k0 ^ k1)
        x0 = np.asarray(ctr0, np.uint64) + ks[0]
        x1 = np.asarray(ctr1, np.uint64) + ks[1]
        for blk in range(5):
            for r in range(4):
                x0 = x0 + x1
                x1 = _rotl(x1, _ROT[(4 * blk + r) % 8])
                x1 = x1 ^ x0
            inj = blk + 1
            x0 = x0 + ks[inj % 3]
            x1 = x1 + ks[(inj + 1) % 3] + np.uint64(inj)
    return x0, x1


def threefry_gaussian(seed: int, pair_idx):
    """_threefry_gaussian (RandomField.jl:46-55)."""
    pair_idx = np.asarray(pair_idx, np.uint64)
    u1r, u2r = threefry2x64_20(pair_idx, np.zeros_like(pair_idx), np.uint64(seed), np.uint64(0))
    u1 = ((u1r >> np.uint64(11)).astype(np.float64) + 1.0) * 1.1102230246251565e-16
    u2 = ((u2r >> np.uint64(11)).astype(np.float64) + 1.0) * 1.1102230246251565e-16
    r = np.sqrt(-2.0 * np.log(u1))
    th = 2.0 * math.pi * u2
    return r * np.cos(th), r * np.sin(th)


def fill_noise_threefry(n: int, seed: int, chunk: int = 1 << 22) -> np.ndarray:
    """fill_noise_threefry! (RandomField.jl:70-81) as Float32."""
    n3 = n * n * n
    out = np.empty(n3, np.float32)
    npairs = (n3 + 1) // 2
    for p0 in range(0, npairs, chunk):
        p1 = min(npairs, p0 + chunk)
        g1, g2 = threefry_gaussian(seed, np.arange(p0, p1, dtype=np.uint64))
        lo, hi = 2 * p0, min(n3, 2 * p1)
        seg = np.empty(2 * (p1 - p0), np.float64)
        seg[0::2] = g1
        seg[1::2] = g2
        out[lo:hi] = seg[: hi - lo]
    return out.reshape(n, n, n)


class PkTable:
    """load_pk: log-log linear interpolation with linear extrapolation."""

    def __init__(self, k, P):
        self.logk = np.log(np.asarray(k, np.float64))
        self.logP = np.log(np.asarray(P, np.float64))

    def __call__(self, k):
        x = np.log(np.asarray(k, np.float64))
        lk, lp = self.logk, self.logP
        i = np.clip(np.searchsorted(lk, x, side="right") - 1, 0, len(lk) - 2)
        t = (x - lk[i]) / (lk[i + 1] - lk[i])
        return np.exp(lp[i] + t * (lp[i + 1] - lp[i]))


def _freqs(n: int, boxsize: float):
    """FFTW.rfftfreq / fftfreq(n, n dk): index times (n dk) / n, Float64."""
    dk = 2.0 * math.pi / boxsize
    mult = (n * dk) / n
    ix = np.arange(n // 2 + 1, dtype=np.float64)
    iy = np.arange(n, dtype=np.float64)
    iy = np.where(iy < (n + 1) // 2, iy, iy - n)
    kx = ix * mult
    ky = iy * mult
    return kx[None, None, :], ky[None, :, None], ky[:, None, None], dk


def _rfft(a):
    import scipy.fft as sf
    return sf.rfftn(a.astype(np.float64), axes=(0, 1, 2), workers=-1).astype(np.complex64)


def _irfft(a, n):
    import scipy.fft as sf
    return sf.irfftn(a.astype(np.complex128), s=(n, n, n), axes=(0, 1, 2), workers=-1).astype(np.float32)


def _mul(c64: np.ndarray, f64) -> np.ndarray:
    """A Float64 factor times a ComplexF32 array, the product rounded back to
    ComplexF32 (Julia's promotion on assignment into the Float32 array)."""
    return (c64.astype(np.complex128) * f64).astype(np.complex64)


def generate_grf(n: int, boxsize: float, seed: int, pk) -> np.ndarray:
    noise = fill_noise_threefry(n, seed)
    nk = _rfft(noise)
    kx, ky, kz, dk = _freqs(n, boxsize)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.sqrt(pk(np.sqrt(k2)) * dk ** 3 * float(n) ** 3)
    amp = np.where(k2 == 0.0, 0.0, amp)
    return _irfft(_mul(nk, amp), n)


def _nyquist_zero(n, shape):
    z = np.zeros(shape, bool)
    z[:, :, -1] = True
    z[:, n // 2, :] = True
    z[n // 2, :, :] = True
    return z


def displacements_1lpt(delta_k, n: int, boxsize: float):
    kx, ky, kz, _ = _freqs(n, boxsize)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    zero = (k2 == 0.0) | _nyquist_zero(n, delta_k.shape)
    k2s = np.where(k2 == 0.0, 1.0, k2)
    out = []
    for kc in (kx, ky, kz):
        f = np.where(zero, 0.0, 1j * kc / k2s)
        out.append(_irfft(_mul(delta_k, f), n))
    return tuple(out)


def displacements_2lpt(delta_k, n: int, boxsize: float):
    kx, ky, kz, _ = _freqs(n, boxsize)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2s = np.where(k2 == 0.0, 1.0, k2)
    kk = (kx, ky, kz)
    phi = {}
    for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)):
        f = np.where(k2 == 0.0, 0.0, -kk[a] * kk[b] / k2s)
        phi[(a, b)] = _irfft(_mul(delta_k, f), n)
    p = phi
    src = p[(0, 0)] * p[(1, 1)]
    src = src - p[(0, 1)] * p[(0, 1)]
    src = src + p[(0, 0)] * p[(2, 2)]
    src = src - p[(0, 2)] * p[(0, 2)]
    src = src + p[(1, 1)] * p[(2, 2)]
    src = src - p[(1, 2)] * p[(1, 2)]
    sk = _rfft(src.astype(np.float32))
    zero = (k2 == 0.0) | _nyquist_zero(n, sk.shape)
    out = []
    for kc in kk:
        f = np.where(zero, 0.0, -1j * kc / k2s)
        out.append(_irfft(_mul(sk, f), n))
    return tuple(out)


def compute_laplacian(delta_k, n: int, boxsize: float) -> np.ndarray:
    dk = np.float32(2.0 * math.pi / boxsize)
    nyq = n // 2
    ix = np.arange(n // 2 + 1)
    iy = np.arange(n)
    iy = np.where(iy <= nyq, iy, iy - n)
    kx = (ix.astype(np.float32) * dk)[None, None, :]
    ky = (iy.astype(np.float32) * dk)[None, :, None]
    kz = (iy.astype(np.float32) * dk)[:, None, None]
    k2 = (kx * kx + ky * ky) + kz * kz
    lk = (delta_k * k2.astype(np.complex64)).astype(np.complex64)
    return (_irfft(lk, n) / np.float32(n) ** 3).astype(np.float32)


class TileFields:
    def __init__(self, n, boxsize, delta, delta_k, eta, eta2, lapd):
        self.n, self.boxsize = n, boxsize
        self.alatt = boxsize / n
        self.delta, self.delta_k = delta, delta_k
        self.eta, self.eta2, self.lapd = eta, eta2, lapd


def make_tile(n: int, boxsize: float, seed: int, pk, ilpt: int = 2, want_lapd: bool = True) -> TileFields:
    """The inputs run_tile builds for a tile (Pipeline.jl:110-206)."""
    delta = generate_grf(n, boxsize, seed, pk)
    delta_k = _rfft(delta)
    eta = displacements_1lpt(delta_k, n, boxsize)
    eta2 = displacements_2lpt(delta_k, n, boxsize) if ilpt >= 2 else None
    lapd = compute_laplacian(delta_k, n, boxsize) if want_lapd else None
    return TileFields(n, boxsize, delta, delta_k, eta, eta2, lapd)


# ----------------------------------------------------------------------------
# Multi-resolution split (run_multitile_split, src/MultiResolution.jl:523-850)
# ----------------------------------------------------------------------------
# A global N^3 Threefry field (N = nsub ntile + 2 nbuff, nsub = nmesh -
# 2 nbuff) is split at the white-noise level: a coarse M^3 grid of block
# sums carries the long modes (periodic convolution on the coarse grid,
# Catmull-Rom interpolated onto each tile), and each tile's residual noise
# (fine minus its coarse block mean) is convolved in isolation on a zero-padded
# (2 nmesh)^3 grid.  Tile indices it, jt, kt are 1-based as in the reference.


class SplitGeom:
    """Geometry of run_multitile_split (MultiResolution.jl:546-553, 612-613)."""

    def __init__(self, nmesh: int, nbuff: int, ntile: int, M: int, alatt: float):
        self.nmesh, self.nbuff, self.ntile, self.M, self.alatt = nmesh, nbuff, ntile, M, alatt
        self.nsub = nmesh - 2 * nbuff
        self.N = self.nsub * ntile + 2 * nbuff
        if self.nsub < 1 or self.N % M:
            raise ValueError(f"N={self.N} must be divisible by M={M} (nsub={self.nsub})")
        self.block = self.N // M
        self.boxsize_full = self.N * alatt
        self.boxsize_local = nmesh * alatt
        self.dcore_box = self.nsub * alatt

    def tile_center(self, it, jt, kt):
        """tile_center (src/MultiTile.jl:79-86)."""
        cen = (self.ntile + 1) / 2
        return ((it - cen) * self.dcore_box, (jt - cen) * self.dcore_box, (kt - cen) * self.dcore_box)

    def tile_index(self, it, jt, kt, nshell: int = 0):
        """1-based global indices gi, gj, gk of the tile's cells along each
        axis (mod1 wrap), MultiResolution.jl:221-229."""
        out = []
        for t in (it, jt, kt):
            o = (t - 1) * self.nsub + 1 - self.nbuff - nshell
            li = np.arange(1, self.nmesh + 2 * nshell + 1)
            out.append((o + li - 2) % self.N + 1)
        return out


def _noise_at(seed: int, gi, gj, gk, N: int) -> np.ndarray:
    """The global Threefry normal at 1-based (gi, gj, gk), Float64 (before
    any Float32 rounding), [k, j, i] broadcast."""
    idx = (gi[None, None, :].astype(np.int64) + (gj[None, :, None].astype(np.int64) - 1) * N
           + (gk[:, None, None].astype(np.int64) - 1) * N * N)
    pair = ((idx - 1) >> 1).astype(np.uint64)
    g1, g2 = threefry_gaussian(seed, pair)
    return np.where((idx - 1) % 2 == 0, g1, g2)


def downsample_noise(N: int, M: int, seed: int) -> np.ndarray:
    """_downsample_noise (MultiResolution.jl:58-83): each coarse cell is the
    Float64 sum of its block^3 fine normals over sqrt(block^3), stored as
    Float32.  (The sum's association differs from the reference's
    sequential loop by Float64 rounding only.)"""
    b = N // M
    out = np.empty((M, M, M), np.float64)
    gi = np.arange(1, N + 1)
    for K in range(M):
        gk = np.arange(K * b + 1, (K + 1) * b + 1)
        v = _noise_at(seed, gi, gi, gk, N)                       # (b, N, N)
        out[K] = v.reshape(b, M, b, M, b).sum(axis=(0, 2, 4))
    return (out / math.sqrt(float(b) * float(b) * float(b))).astype(np.float32)


def _freqs_n(n: int, dk: float):
    """rfftfreq / fftfreq(n, n dk) as Float64 [k, j, i] broadcasts."""
    mult = (n * dk) / n
    ix = np.arange(n // 2 + 1, dtype=np.float64)
    iy = np.arange(n, dtype=np.float64)
    iy = np.where(iy < (n + 1) // 2, iy, iy - n)
    return (ix * mult)[None, None, :], (iy * mult)[None, :, None], (iy * mult)[:, None, None]


def _amp_kernel(n: int, dk: float, pk, kernel: int, dim: int):
    """amp(k) = sqrt(P(k) dk^3 n^3) times the transfer kernel (0: 1; 1: i k_d
    / k^2), k = 0 zeroed, as the complex factor Julia forms
    (MultiResolution.jl:262-276, 288-307, 311-314)."""
    kx, ky, kz = _freqs_n(n, dk)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.sqrt(pk(np.sqrt(k2)) * (dk * dk * dk) * float(n) ** 3)
        if kernel == 0:
            f = amp + 0j
        else:
            kd = (kx, ky, kz)[dim]
            f = 1j * (amp * (kd / k2))
    return np.where(k2 == 0.0, 0.0, f)


def periodic_convolve(noise, boxsize: float, pk, kernel: int = 0, dim: int = 0) -> np.ndarray:
    """rfft -> _periodic_convolve! -> irfft (MultiResolution.jl:288-307) on
    a periodic n^3 grid of side boxsize."""
    n = noise.shape[0]
    return _irfft(_mul(_rfft(noise), _amp_kernel(n, 2.0 * math.pi / boxsize, pk, kernel, dim)), n)


def coarse_fields(g: SplitGeom, seed: int, pk):
    """Phase 1a (MultiResolution.jl:617-631): the coarse noise and its
    periodic convolutions, delta_coarse and psi_coarse[1..3] (Float32)."""
    cn = downsample_noise(g.N, g.M, seed)
    delta = periodic_convolve(cn, g.boxsize_full, pk, 0)
    psi = tuple(periodic_convolve(cn, g.boxsize_full, pk, 1, d) for d in range(3))
    return cn, delta, psi


def extended_residual(g: SplitGeom, it, jt, kt, seed: int, coarse_noise, nshell: int = 0) -> np.ndarray:
    """_generate_extended_residual (MultiResolution.jl:197-236), Float32:
    fine normal (rounded to Float32) minus coarse[ci, cj, ck] times
    Float32(1 / sqrt(block^3))."""
    gi, gj, gk = g.tile_index(it, jt, kt, nshell)
    fine = _noise_at(seed, gi, gj, gk, g.N).astype(np.float32)
    inv = np.float32(1.0 / math.sqrt(float(g.block) ** 3))
    ci, cj, ck = (gi - 1) // g.block, (gj - 1) // g.block, (gk - 1) // g.block
    mean = coarse_noise[ck[:, None, None], cj[None, :, None], ci[None, None, :]] * inv
    return (fine - mean).astype(np.float32)


def isolated_convolve(noise, pk, boxsize_local: float, n: int, kernel: int = 0, dim: int = 0,
                      nshell: int = 0) -> np.ndarray:
    """_isolated_convolve (MultiResolution.jl:246-283): zero-pad to twice the
    input side, rfft, times amp(k) [x kernel] with k on the padded box, k = 0
    zeroed, normalised irfft, the tile region cut out."""
    n_in = noise.shape[0]
    dx = boxsize_local / n
    n2 = 2 * n_in
    padded = np.zeros((n2, n2, n2), np.float32)
    padded[:n_in, :n_in, :n_in] = noise
    pk_ = _rfft(padded)
    dk = 2.0 * math.pi / (n2 * dx)
    out = _irfft(_mul(pk_, _amp_kernel(n2, dk, pk, kernel, dim)), n2)
    s = slice(nshell, nshell + n)
    return np.ascontiguousarray(out[s, s, s])


def catmull_rom(t):
    """_catmull_rom (MultiResolution.jl:156-164), Float32."""
    t = np.float32(t) if np.isscalar(t) else t.astype(np.float32)
    h, o, f = np.float32(0.5), np.float32(1.0), np.float32(1.5)
    t2 = t * t
    t3 = t2 * t
    w0 = (-h * t3 + t2) - h * t
    w1 = (f * t3 - np.float32(2.5) * t2) + o
    w2 = (-f * t3 + np.float32(2.0) * t2) + h * t
    w3 = h * t3 - h * t2
    return w0, w1, w2, w3


def interpolate_to_tile(coarse, g: SplitGeom, it, jt, kt) -> np.ndarray:
    """_interpolate_to_tile + _tricubic (MultiResolution.jl:95-122, 167-187):
    cell-centred coarse coordinate f = (g - 0.5) / block + 0.5, Catmull-Rom
    weights of its fraction, 64 taps with periodic wrap, accumulated in
    Float32 in the order k, j, i."""
    M = g.M
    gi, gj, gk = g.tile_index(it, jt, kt)
    base, w = [], []
    for ga in (gi, gj, gk):
        f = (ga - 0.5) / g.block + 0.5
        i0 = np.floor(f).astype(np.int64)
        base.append(i0)
        w.append(catmull_rom((f - i0).astype(np.float32)))
    (bx, by, bz), (wx, wy, wz) = base, w
    val = np.zeros((len(gk), len(gj), len(gi)), np.float32)
    for c in range(4):
        kk = (bz + c - 1 - 1) % M
        for b in range(4):
            jj = (by + b - 1 - 1) % M
            for a in range(4):
                ii = (bx + a - 1 - 1) % M
                wgt = (wx[a][None, None, :] * wy[b][None, :, None]) * wz[c][:, None, None]
                val = val + wgt * coarse[kk[:, None, None], jj[None, :, None], ii[None, None, :]]
    return val


def _apply_kernel(arr_k, n: int, boxsize: float, kern):
    """_apply_kernel_inplace! (MultiResolution.jl:1239-1255): k = 0 and the
    Nyquist planes zeroed, every other mode times kern(kx, ky, kz, k2)
    (Float64), rounded back to ComplexF32."""
    dk = 2.0 * math.pi / boxsize
    kx, ky, kz = _freqs_n(n, dk)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    zero = (k2 == 0.0) | _nyquist_zero(n, arr_k.shape)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = kern(kx, ky, kz, np.where(k2 == 0.0, 1.0, k2))
    f = np.where(zero, 0.0, f)
    return _mul(arr_k, f)


def split_2lpt(delta_tile, n: int, boxsize_local: float):
    """The CPU branch of process_tile!'s 2LPT (MultiResolution.jl:811-831):
    src2 = delta^2 / 2 - sum_d phi_dd^2 / 2 - sum_{a<b} phi_ab^2 in Float32
    in that order, then psi2_d = irfft(-i k_d / k^2 rfft(src2))."""
    dk_ = _rfft(delta_tile)
    h = np.float32(0.5)
    src = np.zeros_like(delta_tile) + (delta_tile * delta_tile) * h
    for d in range(3):
        phi = _irfft(_apply_kernel(dk_, n, boxsize_local,
                                   lambda kx, ky, kz, k2, d=d: -(kx, ky, kz)[d] * (kx, ky, kz)[d] / k2), n)
        src = src - (phi * phi) * h
    for a, b in ((0, 1), (0, 2), (1, 2)):
        phi = _irfft(_apply_kernel(dk_, n, boxsize_local,
                                   lambda kx, ky, kz, k2, a=a, b=b: -(kx, ky, kz)[a] * (kx, ky, kz)[b] / k2), n)
        src = src - phi * phi
    sk = _rfft(src.astype(np.float32))
    return tuple(_irfft(_apply_kernel(sk, n, boxsize_local,
                                      lambda kx, ky, kz, k2, d=d: -1j * ((kx, ky, kz)[d] / k2)), n)
                 for d in range(3))


def split_laplacian(delta_tile, n: int, boxsize_local: float) -> np.ndarray:
    """process_tile!'s k^2 field (MultiResolution.jl:843-847): irfft of
    k^2 delta_k with k = 0 and the Nyquist planes zeroed (one
    normalisation)."""
    return _irfft(_apply_kernel(_rfft(delta_tile), n, boxsize_local, lambda kx, ky, kz, k2: k2 + 0.0 * kx), n)


def split_tile(g: SplitGeom, it, jt, kt, seed: int, pk, coarse=None, ilpt: int = 2, want_lapd: bool = True):
    """One tile's fields on the CPU path of process_tile!
    (MultiResolution.jl:719-850): residual -> isolated delta and psi, plus
    the interpolated coarse fields; 2LPT and k^2 from the combined delta."""
    if coarse is None:
        coarse = coarse_fields(g, seed, pk)
    cn, dc, pc = coarse
    n = g.nmesh
    res = extended_residual(g, it, jt, kt, seed, cn)
    delta = (isolated_convolve(res, pk, g.boxsize_local, n, 0) + interpolate_to_tile(dc, g, it, jt, kt))
    eta = tuple((isolated_convolve(res, pk, g.boxsize_local, n, 1, d) + interpolate_to_tile(pc[d], g, it, jt, kt))
                .astype(np.float32) for d in range(3))
    delta = delta.astype(np.float32)
    eta2 = split_2lpt(delta, n, g.boxsize_local) if ilpt >= 2 else None
    lapd = split_laplacian(delta, n, g.boxsize_local) if want_lapd else None
    return TileFields(n, g.boxsize_local, delta, None, eta, eta2, lapd)
