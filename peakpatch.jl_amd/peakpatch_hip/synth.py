"""Seeded synthetic inputs for the halo-finding path.

The reference builds its inputs in ``src/InitialConditions`` (Threefry white
noise, P(k) convolution, LPT displacements) and ``src/Pipeline.jl:35-51``
(the k^2-weighted "Laplacian" field).  This module restates those producers
so that the filter bank, the peak stencil and the shell analysis can be fed
fields of the configured shape:

* ``threefry2x64_20`` / ``fill_noise_threefry``: Random123's Threefry-2x64
  with 20 rounds, keyed (seed, 0) on counter (pair_idx, 0), Box-Muller on the
  two 53-bit uniforms (RandomField.jl:46-79).
* ``eh98_nowiggle_pk``: Eisenstein & Hu (1998, ApJ 496, 605) eqs. 26-31
  no-wiggle transfer function, n_s tilt, sigma_8 normalised with the real-space
  top-hat (MassFunction.jl:29-42 convention), then divided by (2 pi)^3 so that
  <delta^2> = int P d^3k, the field-generator convention of
  RandomField.jl:136-175 (the reference reads P(k) from a file in that convention).
* ``generate_grf`` (RandomField.jl:123-178), ``displacements_1lpt`` /
  ``displacements_2lpt`` (LPT.jl:21-191), ``compute_laplacian``
  (Pipeline.jl:35-51).

Arrays use the reference's memory layout: a Julia ``Array{Float32,3}`` of size
(n, n, n) is x-fastest, i.e. a C-order numpy array of shape (n, n, n) indexed
[k, j, i].  FFTs run on scipy's pocketfft (CPU) or, for the large bench
tiles, on torch's device FFT when a torch device is given (input plumbing
only: the hot path never calls these).
"""
from __future__ import annotations

import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_ROT = (16, 42, 12, 31, 16, 32, 24, 21)
_PARITY = np.uint64(0x1BD11BDAA9FC1A22)


def _rotl(x: np.ndarray, r: int) -> np.ndarray:
    return (x << np.uint64(r)) | (x >> np.uint64(64 - r))


def threefry2x64_20(ctr0, ctr1, key0, key1):
    """Threefry-2x64-20 block function (Salmon et al. 2011, Random123).
    Inputs are uint64 scalars or arrays; returns (x0, x1)."""
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


def threefry_gaussian(seed: int, pair_idx: np.ndarray):
    """_threefry_gaussian (RandomField.jl:46-55)."""
    pair_idx = np.asarray(pair_idx, np.uint64)
    u1r, u2r = threefry2x64_20(pair_idx, np.zeros_like(pair_idx), np.uint64(seed), np.uint64(0))
    u1 = ((u1r >> np.uint64(11)).astype(np.float64) + 1.0) * 1.1102230246251565e-16
    u2 = ((u2r >> np.uint64(11)).astype(np.float64) + 1.0) * 1.1102230246251565e-16
    r = np.sqrt(-2.0 * np.log(u1))
    th = 2.0 * math.pi * u2
    return r * np.cos(th), r * np.sin(th)


def fill_noise_threefry(n: int, seed: int, dtype=np.float32, chunk: int = 1 << 22) -> np.ndarray:
    """fill_noise_threefry! (RandomField.jl:70-81): cell idx (1-based,
    x-fastest) takes g1 of pair (idx-1)>>1 when idx-1 is even, else g2."""
    n3 = n * n * n
    out = np.empty(n3, dtype)
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


def _kgrid(n: int, boxsize: float):
    dk = 2.0 * math.pi / boxsize
    kx = dk * np.fft.rfftfreq(n, 1.0 / n)
    ky = dk * np.fft.fftfreq(n, 1.0 / n)
    return kx[None, None, :], ky[None, :, None], ky[:, None, None]


def _fft():
    import scipy.fft as sf
    return sf


def generate_grf(n: int, boxsize: float, seed: int, pk, workers: int = -1) -> tuple[np.ndarray, np.ndarray]:
    """generate_grf (RandomField.jl:123-178): Threefry noise -> rfft ->
    * sqrt(P(k) dk^3 n^3), DC zeroed -> irfft.  Returns (delta float32,
    delta_k complex64 = rfft(delta))."""
    sf = _fft()
    noise = fill_noise_threefry(n, seed)
    nk = sf.rfftn(noise, workers=workers)                         # complex64
    kx, ky, kz = _kgrid(n, boxsize)
    k2 = kx * kx + ky * ky + kz * kz
    dk = 2.0 * math.pi / boxsize
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.sqrt(pk(np.sqrt(k2)) * dk ** 3 * float(n) ** 3)
    amp[0, 0, 0] = 0.0
    nk = (nk * amp).astype(np.complex64)
    delta = sf.irfftn(nk, s=(n, n, n), workers=workers).astype(np.float32)
    delta_k = sf.rfftn(delta, workers=workers).astype(np.complex64)
    return delta, delta_k


def displacements_1lpt(delta_k: np.ndarray, n: int, boxsize: float, workers: int = -1):
    """displacements_1lpt (LPT.jl:21-71): psi_i = i k_i / k^2 delta_k,
    k=0 and Nyquist planes zeroed."""
    sf = _fft()
    kx, ky, kz = _kgrid(n, boxsize)
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    keep = np.ones(delta_k.shape, bool)
    keep[:, :, -1] = False
    keep[:, n // 2, :] = False
    keep[n // 2, :, :] = False
    keep[0, 0, 0] = False
    out = []
    for kc in (kx, ky, kz):
        pk = np.where(keep, 1j * kc / k2 * delta_k, 0).astype(np.complex64)
        out.append(sf.irfftn(pk, s=(n, n, n), workers=workers).astype(np.float32))
    return tuple(out)


def displacements_2lpt(delta_k: np.ndarray, n: int, boxsize: float, workers: int = -1):
    """displacements_2lpt (LPT.jl:86-191)."""
    sf = _fft()
    kx, ky, kz = _kgrid(n, boxsize)
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    kk = (kx, ky, kz)
    phi = {}
    for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)):
        pk = (-kk[a] * kk[b] / k2 * delta_k).astype(np.complex64)
        pk[0, 0, 0] = 0
        phi[(a, b)] = sf.irfftn(pk, s=(n, n, n), workers=workers).astype(np.float32)
    src = (phi[(0, 0)] * phi[(1, 1)] - phi[(0, 1)] ** 2 + phi[(0, 0)] * phi[(2, 2)] - phi[(0, 2)] ** 2
           + phi[(1, 1)] * phi[(2, 2)] - phi[(1, 2)] ** 2)
    del phi
    sk = sf.rfftn(src, workers=workers)
    keep = np.ones(sk.shape, bool)
    keep[:, :, -1] = False
    keep[:, n // 2, :] = False
    keep[n // 2, :, :] = False
    keep[0, 0, 0] = False
    out = []
    for kc in kk:
        pk = np.where(keep, -1j * kc / k2 * sk, 0).astype(np.complex64)
        out.append(sf.irfftn(pk, s=(n, n, n), workers=workers).astype(np.float32))
    return tuple(out)


def compute_laplacian(delta_k: np.ndarray, n: int, boxsize: float, workers: int = -1) -> np.ndarray:
    """_compute_laplacian (Pipeline.jl:35-51): irfft(k^2 delta_k) / n^3 in
    Float32 (the reference divides the already-normalised irfft by n^3 again)."""
    sf = _fft()
    kx, ky, kz = _kgrid(n, boxsize)
    k2 = (kx * kx + ky * ky + kz * kz).astype(np.float32)
    lk = (k2 * delta_k).astype(np.complex64)
    return (sf.irfftn(lk, s=(n, n, n), workers=workers) / np.float32(n) ** 3).astype(np.float32)


def filter_bank(nfilt: int, rmin: float, rmax: float) -> np.ndarray:
    """Log-spaced top-hat radii, returned in the driver's order: descending
    (Pipeline.jl:81-82)."""
    r = np.exp(np.linspace(math.log(rmin), math.log(rmax), nfilt))
    return np.sort(r)[::-1].copy()


class TileInputs:
    """The fields one tile of the path consumes (PeakGrid, RadialShell.jl:45-56)
    plus delta_k for the filter bank."""

    def __init__(self, n, boxsize, delta, delta_k, eta, eta2, lapd):
        self.n, self.boxsize = n, boxsize
        self.alatt = boxsize / n
        self.delta, self.delta_k = delta, delta_k
        self.eta, self.eta2, self.lapd = eta, eta2, lapd


def make_tile(n: int, boxsize: float, seed: int = 42, ilpt: int = 2, want_lapd: bool = True,
              pk=None, workers: int = -1) -> TileInputs:
    pk = pk or EH98NoWiggle()
    delta, delta_k = generate_grf(n, boxsize, seed, pk, workers)
    eta = displacements_1lpt(delta_k, n, boxsize, workers)
    eta2 = displacements_2lpt(delta_k, n, boxsize, workers) if ilpt >= 2 else None
    lapd = compute_laplacian(delta_k, n, boxsize, workers) if want_lapd else None
    return TileInputs(n, boxsize, delta, delta_k, eta, eta2, lapd)


# ---------------------------------------------------------------------------
# Device (torch) version of the same producers, for bench-size tiles.  Same
# formulas as above; Threefry runs on int64 tensors (two's-complement adds,
# masked arithmetic shifts give the logical rotations).
# ---------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def _t_rotl(x, r: int):
    return (x << r) | ((x >> (64 - r)) & ((1 << r) - 1))


def threefry2x64_20_torch(ctr0, key0: int):
    """Threefry-2x64-20 of counters (ctr0, 0) under key (key0, 0)."""
    torch = _torch()
    par = 0x1BD11BDAA9FC1A22
    ks = (key0, 0, par ^ key0)
    x0 = ctr0 + ks[0]
    x1 = torch.zeros_like(ctr0) + ks[1]
    for blk in range(5):
        for r in range(4):
            x0 = x0 + x1
            x1 = _t_rotl(x1, _ROT[(4 * blk + r) % 8])
            x1 = x1 ^ x0
        inj = blk + 1
        x0 = x0 + ks[inj % 3]
        x1 = x1 + (ks[(inj + 1) % 3] + inj)
    return x0, x1


def fill_noise_threefry_torch(n: int, seed: int, device, chunk: int = 1 << 24):
    torch = _torch()
    n3 = n * n * n
    npairs = (n3 + 1) // 2
    out = torch.empty(2 * npairs, dtype=torch.float32, device=device)
    m53 = (1 << 53) - 1
    for p0 in range(0, npairs, chunk):
        p1 = min(npairs, p0 + chunk)
        ctr = torch.arange(p0, p1, dtype=torch.int64, device=device)
        u1r, u2r = threefry2x64_20_torch(ctr, int(seed))
        u1 = (((u1r >> 11) & m53).to(torch.float64) + 1.0) * 1.1102230246251565e-16
        u2 = (((u2r >> 11) & m53).to(torch.float64) + 1.0) * 1.1102230246251565e-16
        r = torch.sqrt(-2.0 * torch.log(u1))
        th = 2.0 * math.pi * u2
        seg = torch.stack([r * torch.cos(th), r * torch.sin(th)], dim=1).reshape(-1)
        out[2 * p0:2 * p1] = seg.to(torch.float32)
    return out[:n3].reshape(n, n, n)


def _kgrid_torch(n, boxsize, device):
    torch = _torch()
    dk = 2.0 * math.pi / boxsize
    kx = dk * torch.fft.rfftfreq(n, 1.0 / n, dtype=torch.float64, device=device)
    ky = dk * torch.fft.fftfreq(n, 1.0 / n, dtype=torch.float64, device=device)
    return kx[None, None, :], ky[None, :, None], ky[:, None, None]


def make_tile_torch(n: int, boxsize: float, seed: int = 42, ilpt: int = 2, want_lapd: bool = True,
                    pk=None, device="cuda", ctx=None):
    """make_tile on a torch device; returns TileInputs holding torch tensors
    (delta_k stays on the device as complex64).  With a Context the white
    noise comes from the library's Threefry kernel on its stream."""
    torch = _torch()
    pk = pk or EH98NoWiggle()
    noise = ctx.fill_noise(n, seed) if ctx is not None else fill_noise_threefry_torch(n, seed, device)
    kx, ky, kz = _kgrid_torch(n, boxsize, device)
    k2 = kx * kx + ky * ky + kz * kz
    dk = 2.0 * math.pi / boxsize
    kmag = torch.sqrt(k2)
    # P(k) evaluated on the host over the distinct |k| would be exact; the
    # device evaluates the same closed form in float64.
    amp = torch.sqrt(_pk_torch(pk, kmag) * dk ** 3 * float(n) ** 3)
    amp[0, 0, 0] = 0.0
    nk = torch.fft.rfftn(noise)
    del noise
    nk = nk * amp.to(torch.float32)
    delta = torch.fft.irfftn(nk, s=(n, n, n)).to(torch.float32).contiguous()
    del nk, amp
    delta_k = torch.fft.rfftn(delta)
    k2s = k2.clone()
    k2s[0, 0, 0] = 1.0
    keep = torch.ones(delta_k.shape, dtype=torch.bool, device=device)
    keep[:, :, -1] = False
    keep[:, n // 2, :] = False
    keep[n // 2, :, :] = False
    keep[0, 0, 0] = False
    eta = []
    for kc in (kx, ky, kz):
        m = (kc / k2s).to(torch.float32)
        eta.append(torch.fft.irfftn(torch.where(keep, 1j * m * delta_k, 0), s=(n, n, n)).contiguous())
    eta2 = None
    if ilpt >= 2:
        kk = (kx, ky, kz)
        phi = {}
        for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)):
            m = (-kk[a] * kk[b] / k2s).to(torch.float32)
            t = m * delta_k
            t[0, 0, 0] = 0
            phi[(a, b)] = torch.fft.irfftn(t, s=(n, n, n))
        src = (phi[(0, 0)] * phi[(1, 1)] - phi[(0, 1)] ** 2 + phi[(0, 0)] * phi[(2, 2)] - phi[(0, 2)] ** 2
               + phi[(1, 1)] * phi[(2, 2)] - phi[(1, 2)] ** 2)
        del phi
        sk = torch.fft.rfftn(src)
        del src
        eta2 = []
        for kc in kk:
            m = (kc / k2s).to(torch.float32)
            eta2.append(torch.fft.irfftn(torch.where(keep, -1j * m * sk, 0), s=(n, n, n)).contiguous())
        del sk
    lapd = None
    if want_lapd:
        lapd = (torch.fft.irfftn(k2.to(torch.float32) * delta_k, s=(n, n, n)) / float(n) ** 3).contiguous()
    return TileInputs(n, boxsize, delta, delta_k, tuple(eta), None if eta2 is None else tuple(eta2), lapd)


def _pk_torch(pk: EH98NoWiggle, k):
    torch = _torch()
    ks = k * pk.h * pk.s
    gamma = pk.Om * pk.h * (pk.alpha + (1.0 - pk.alpha) / (1.0 + (0.43 * ks) ** 4))
    q = k * pk.theta ** 2 / gamma
    L0 = torch.log(2.0 * math.e + 1.8 * q)
    C0 = 14.2 + 731.0 / (1.0 + 62.5 * q)
    T = L0 / (L0 + C0 * q * q)
    kk = torch.where(k > 0, k, torch.ones_like(k))
    return torch.where(k > 0, pk.A * kk ** pk.ns * T * T / (2.0 * math.pi) ** 3, torch.zeros_like(k))
