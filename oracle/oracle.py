"""ctypes front-end for the FP64 CPU oracle (oracle/pp_oracle.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this module, and only to check or
time the product against the reference algorithm.  The product package never
imports it.

Each wrapper mirrors one reference function (file:line cited in the C source).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")


def build() -> str:
    """Compile liboracle.so with gcc (Makefile beside this file)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    return _SO


def _load():
    if not os.path.exists(_SO):
        build()
    return C.CDLL(_SO)


_lib = _load()
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C")


class OrcTable(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("X1", C.c_double), ("X2", C.c_double), ("Y1", C.c_double),
                ("Y2", C.c_double), ("Z1", C.c_double), ("Z2", C.c_double),
                ("data", C.c_void_p)]


class OrcGrid(C.Structure):
    _fields_ = [("delta", C.c_void_p), ("etax", C.c_void_p), ("etay", C.c_void_p),
                ("etaz", C.c_void_p), ("eta2x", C.c_void_p), ("eta2y", C.c_void_p),
                ("eta2z", C.c_void_p), ("mask", C.c_void_p), ("lapd", C.c_void_p),
                ("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int)]


class OrcShells(C.Structure):
    _fields_ = [("di", C.c_void_p), ("dj", C.c_void_p), ("dk", C.c_void_p),
                ("r2", C.c_void_p), ("ncells", C.c_int64)]


class OrcResult(C.Structure):
    _fields_ = [("RTHL", C.c_double), ("Srb", C.c_double), ("Sbar", C.c_double * 3),
                ("Sbar2", C.c_double * 3), ("strain", C.c_double * 9),
                ("gradpk", C.c_double * 3), ("gradpkrf", C.c_double * 3),
                ("Fbarx", C.c_double), ("e_v", C.c_double), ("p_v", C.c_double),
                ("zvir_half", C.c_double), ("gradpkf", C.c_double * 3), ("d2F", C.c_double)]


class OrcProfile(C.Structure):
    _fields_ = [("m", C.c_int), ("cap", C.c_int), ("nshell", C.c_void_p),
                ("Fbar", C.c_void_p), ("Fsum", C.c_void_p), ("m0_final", C.c_int)]


RESULT_DTYPE = np.dtype([("RTHL", "f8"), ("Srb", "f8"), ("Sbar", "f8", 3), ("Sbar2", "f8", 3),
                         ("strain", "f8", 9), ("gradpk", "f8", 3), ("gradpkrf", "f8", 3),
                         ("Fbarx", "f8"), ("e_v", "f8"), ("p_v", "f8"), ("zvir_half", "f8"),
                         ("gradpkf", "f8", 3), ("d2F", "f8")])
assert RESULT_DTYPE.itemsize == C.sizeof(OrcResult)

DUMP_NAMES = ("none", "premask", "no_fcrit", "m0_one", "no_collapse", "rthl_neg", "nonfinite")

_lib.orc_hRinteg.restype = C.c_double
_lib.orc_hRinteg.argtypes = [C.c_double, C.c_double]
_lib.orc_atab4.argtypes = [_f64p]
_lib.orc_precompute_shells.restype = C.c_int64
_lib.orc_precompute_shells.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
_lib.orc_normalize_strain.argtypes = [_f64p, C.c_double]
_lib.orc_get_evals.argtypes = [_f64p, _f64p]
_lib.orc_interp_c.restype = C.c_double
_lib.orc_interp_c.argtypes = [_f32p, C.c_int, C.c_int, C.c_int] + [C.c_double] * 9
_lib.orc_tophat_window.restype = C.c_double
_lib.orc_tophat_window.argtypes = [C.c_double]
_lib.orc_gaussian_window.restype = C.c_double
_lib.orc_gaussian_window.argtypes = [C.c_double]
_lib.orc_gaussian_window_fortran.restype = C.c_double
_lib.orc_gaussian_window_fortran.argtypes = [C.c_double]
_lib.orc_window_multiply.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int]
_lib.orc_find_peaks.restype = C.c_int64
_lib.orc_find_peaks.argtypes = [_f32p, _i8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                C.c_int64, _i32p, _i32p, _i32p, _f32p]
_lib.orc_dlin_size.restype = C.c_size_t
_lib.orc_dlinear_tables.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double]
_lib.orc_dlinear_ab.restype = C.c_double
_lib.orc_dlinear_ab.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
_lib.orc_fsc_of_z.restype = C.c_double
_lib.orc_fsc_of_z.argtypes = [C.c_void_p, C.c_double]
_lib.orc_chi.restype = C.c_double
_lib.orc_chi.argtypes = [C.c_double] * 4
_lib.orc_chi_to_z.restype = C.c_double
_lib.orc_chi_to_z.argtypes = [_f64p, _f64p, C.c_int, C.c_double, C.c_double]
_lib.orc_build_chi_to_z.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _f64p, _f64p]
_lib.orc_analyse_peak.restype = C.c_int
_lib.orc_analyse_peak.argtypes = [C.POINTER(OrcGrid), C.c_int64, C.c_double, C.c_int64, C.c_double,
                                  C.c_double, C.POINTER(OrcTable), C.POINTER(OrcShells), C.c_int,
                                  C.c_double, C.c_double, C.POINTER(OrcResult), C.POINTER(OrcProfile)]
_lib.orc_analyse_peaks.restype = C.c_int
_lib.orc_analyse_peaks.argtypes = [C.POINTER(OrcGrid), C.c_int64, _i64p, _i64p, _f64p, _f64p, _f64p,
                                   C.c_double, C.POINTER(OrcTable), C.POINTER(OrcShells), C.c_int,
                                   C.c_double, C.c_int, C.c_void_p, _i32p]


def hRinteg(u0: float, u12: float) -> float:
    return _lib.orc_hRinteg(u0, u12)


def atab4() -> np.ndarray:
    a = np.zeros(1101, np.float64)
    _lib.orc_atab4(a)
    return a


class Shells:
    """precompute_shells(rmax) as SoA int32 arrays (RadialShell.jl:160-175)."""

    def __init__(self, rmax: int):
        n = _lib.orc_precompute_shells(rmax, None, None, None, None, 0)
        self.rmax = rmax
        self.di = np.zeros(n, np.int32)
        self.dj = np.zeros(n, np.int32)
        self.dk = np.zeros(n, np.int32)
        self.r2 = np.zeros(n, np.int32)
        _lib.orc_precompute_shells(rmax, self.di.ctypes.data, self.dj.ctypes.data,
                                   self.dk.ctypes.data, self.r2.ctypes.data, n)
        self.c = OrcShells(self.di.ctypes.data, self.dj.ctypes.data, self.dk.ctypes.data,
                           self.r2.ctypes.data, n)

    def __len__(self):
        return len(self.r2)


def normalize_strain(m: np.ndarray, F: float) -> np.ndarray:
    a = np.asfortranarray(m, dtype=np.float64).ravel(order="F").copy()
    _lib.orc_normalize_strain(a, F)
    return a.reshape(3, 3, order="F")


def get_evals(m: np.ndarray) -> np.ndarray:
    a = np.asarray(m, np.float64).ravel(order="F").copy()
    lam = np.zeros(3)
    _lib.orc_get_evals(a, lam)
    return lam


class Table:
    """Collapse table (CollapseTable.jl:102-145); data x-fastest Float32."""

    def __init__(self, data: np.ndarray, X1, X2, Y1, Y2, Z1, Z2):
        data = np.asarray(data, np.float32)
        self.nx, self.ny, self.nz = data.shape
        self.flat = np.ascontiguousarray(data.ravel(order="F"))
        self.bounds = (float(X1), float(X2), float(Y1), float(Y2), float(Z1), float(Z2))
        self.c = OrcTable(self.nx, self.ny, self.nz, *self.bounds, self.flat.ctypes.data)

    @classmethod
    def read_homeltab(cls, path):
        raw = open(path, "rb").read()
        nx, ny, nz = np.frombuffer(raw[:12], "<i4")
        b = np.frombuffer(raw[12:36], "<f4").astype(np.float64)
        data = np.frombuffer(raw[36:36 + 4 * nx * ny * nz], "<f4").reshape((nx, ny, nz), order="F")
        return cls(data, *b)

    def interp(self, x, y, z) -> float:
        return _lib.orc_interp_c(self.flat, self.nx, self.ny, self.nz, *self.bounds, x, y, z)


def tophat_window(kR):
    return _lib.orc_tophat_window(kR)


def gaussian_window(kR):
    return _lib.orc_gaussian_window(kR)


def gaussian_window_fortran(kR):
    return _lib.orc_gaussian_window_fortran(kR)


def window_multiply(delta_k: np.ndarray, n: int, boxsize: float, Rf: float, wsmooth: int) -> np.ndarray:
    """k-space part of smooth_field: delta_k complex64 shaped (n, n, n//2+1) C-order
    (= (n//2+1, n, n) x-fastest).  Returns a new complex64 array."""
    a = np.ascontiguousarray(delta_k, np.complex64)
    out = np.empty_like(a)
    _lib.orc_window_multiply(a.ctypes.data, out.ctypes.data, n, boxsize, Rf, wsmooth)
    return out


def smooth_field(delta_k: np.ndarray, n: int, boxsize: float, Rf: float, wsmooth: int) -> np.ndarray:
    """smooth_field (Filters.jl:79-106): window multiply (C) + FP64 pocketfft irfft.
    Returns float64 (z, y, x) C-order array == Julia (x, y, z) column-major."""
    wk = window_multiply(delta_k, n, boxsize, Rf, wsmooth)
    return np.fft.irfftn(wk.astype(np.complex128), s=(n, n, n), axes=(0, 1, 2))


def find_peaks(delta: np.ndarray, mask: np.ndarray, nbuff: int, fcrit: float,
               max_peaks: int = 2_000_000):
    """find_peaks (PeakFind.jl:15-78).  delta/mask are (n3, n2, n1) C-order arrays
    (x fastest).  Mutates mask.  Returns 1-based (i, j, k, delta) arrays."""
    n3, n2, n1 = delta.shape
    d = np.ascontiguousarray(delta, np.float32)
    assert mask.dtype == np.int8 and mask.flags.c_contiguous
    cap = min(max_peaks, max(1, (n1 - 2 * nbuff) * (n2 - 2 * nbuff) * (n3 - 2 * nbuff)))
    oi = np.zeros(cap, np.int32); oj = np.zeros(cap, np.int32); ok = np.zeros(cap, np.int32)
    od = np.zeros(cap, np.float32)
    npk = _lib.orc_find_peaks(d, mask, n1, n2, n3, nbuff, np.float32(fcrit), max_peaks, oi, oj, ok, od)
    return oi[:npk], oj[:npk], ok[:npk], od[:npk]


class Dlinear:
    """Dlinear_tables / Dlinear_ab / fsc_of_z (Cosmology.jl:148-293; RadialShell.jl:191-196)."""

    def __init__(self, Om, OB, OL, h):
        self.buf = C.create_string_buffer(_lib.orc_dlin_size())
        _lib.orc_dlinear_tables(self.buf, Om, OB, OL, h)

    def ab(self, a):
        hd = C.c_double(); da = C.c_double()
        d = _lib.orc_dlinear_ab(self.buf, a, C.byref(hd), C.byref(da))
        return d, hd.value, da.value

    def fsc_of_z(self, z):
        return _lib.orc_fsc_of_z(self.buf, z)


def chi(z, Om, OL, h):
    return _lib.orc_chi(z, Om, OL, h)


class ChiToZ:
    """build_chi_to_z / chi_to_z / peak_redshift (Cosmology.jl:333-372)."""

    def __init__(self, Om, OL, h, z_max=10.0, npts=2000):
        self.zs = np.zeros(npts); self.chis = np.zeros(npts)
        self.z_max = z_max; self.npts = npts
        _lib.orc_build_chi_to_z(Om, OL, h, z_max, npts, self.zs, self.chis)

    def __call__(self, chi_val):
        return _lib.orc_chi_to_z(self.zs, self.chis, self.npts, self.z_max, float(chi_val))

    def peak_redshift(self, obs, pk):
        d = np.asarray(pk, np.float64) - np.asarray(obs, np.float64)
        return self(float(np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)))


class Grid:
    """PeakGrid (RadialShell.jl:45-56).  Arrays are (n3, n2, n1) C-order float32."""

    def __init__(self, delta, etax, etay, etaz, eta2=None, mask=None, lapd=None):
        self.keep = [np.ascontiguousarray(a, np.float32) for a in (delta, etax, etay, etaz)]
        e2 = [None, None, None] if eta2 is None else [np.ascontiguousarray(a, np.float32) for a in eta2]
        self.keep += [a for a in e2 if a is not None]
        self.mask = mask
        lp = None if lapd is None else np.ascontiguousarray(lapd, np.float32)
        if lp is not None:
            self.keep.append(lp)
        n3, n2, n1 = delta.shape
        p = lambda a: None if a is None else a.ctypes.data
        self.c = OrcGrid(p(self.keep[0]), p(self.keep[1]), p(self.keep[2]), p(self.keep[3]),
                         p(e2[0]), p(e2[1]), p(e2[2]), p(mask), p(lp), n1, n2, n3)


def analyse_peak(grid: Grid, ipp: int, alatt: float, ir2min: int, ZZon: float, Rf: float,
                 table: Table, shells: Shells, nbuff=0, rmax2rs=0.0, fcrit=1.686,
                 profile_cap: int = 0):
    """analyse_peak (RadialShell.jl:282-891) for one peak; ipp 1-based.
    Returns (result dict, dump status, profile or None)."""
    r = OrcResult()
    prof = None
    pp = None
    if profile_cap > 0:
        prof = {"nshell": np.zeros(profile_cap, np.int64), "Fbar": np.zeros(profile_cap),
                "Fsum": np.zeros(profile_cap)}
        pp = OrcProfile(0, profile_cap, prof["nshell"].ctypes.data, prof["Fbar"].ctypes.data,
                        prof["Fsum"].ctypes.data, 0)
    st = _lib.orc_analyse_peak(C.byref(grid.c), int(ipp), alatt, int(ir2min), ZZon, Rf,
                               C.byref(table.c), C.byref(shells.c), nbuff, rmax2rs, fcrit,
                               C.byref(r), C.byref(pp) if pp is not None else None)
    out = np.frombuffer(bytes(r), RESULT_DTYPE)[0]
    if prof is not None:
        prof["m"] = pp.m
        prof["m0"] = pp.m0_final
    return out, st, prof


def analyse_peaks(grid: Grid, ipp, ir2min, ZZon, Rf, fcrit, alatt, table: Table, shells: Shells,
                  nbuff=0, rmax2rs=0.0, nthreads=0):
    """OpenMP batch of analyse_peak (Pipeline.jl:233-241 threading)."""
    ipp = np.ascontiguousarray(ipp, np.int64)
    n = len(ipp)
    res = np.zeros(n, RESULT_DTYPE)
    st = np.zeros(n, np.int32)
    _lib.orc_analyse_peaks(C.byref(grid.c), n, ipp, np.ascontiguousarray(ir2min, np.int64),
                           np.ascontiguousarray(ZZon, np.float64), np.ascontiguousarray(Rf, np.float64),
                           np.ascontiguousarray(fcrit, np.float64), alatt, C.byref(table.c),
                           C.byref(shells.c), nbuff, rmax2rs, nthreads, res.ctypes.data, st)
    return res, st


# ---- ellipsoidal collapse (EllipsoidalCollapse.jl RK4 path; CollapseTable.jl) ----
_lib.orc_elliptic_rd.restype = C.c_double
_lib.orc_elliptic_rd.argtypes = [C.c_double] * 3
_lib.orc_get_b_2.argtypes = [C.c_double, C.c_double, C.c_double, _f64p]
_lib.orc_evolve_ellipse.restype = C.c_double
_lib.orc_evolve_ellipse.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.POINTER(C.c_int)]
_lib.orc_dfnofa.restype = C.c_double
_lib.orc_dfnofa.argtypes = [C.c_void_p, C.c_double]
_lib.orc_make_collapse_table.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int] + \
    [C.c_double] * 6 + [C.c_void_p]

TABLE_BOUNDS = (math.log10(1.5), math.log10(8.0), 0.0, 0.5, -1.0 + 1e-4, 1.0 - 1e-4)


def elliptic_rd(x, y, z):
    return _lib.orc_elliptic_rd(x, y, z)


def get_b_2(a1, a2, a3):
    b = np.zeros(3)
    _lib.orc_get_b_2(a1, a2, a3, b)
    return b


class Ellipsoid:
    """evolve_ellipse_full with the Fortran-compatible RK4 solver and the
    reference's default parameters, for one cosmology."""

    def __init__(self, Om=0.315, OB=0.049, OL=0.685, h=0.674):
        self.Om, self.OL = Om, OL
        self.dlin = Dlinear(Om, OB, OL, h)

    def zvir(self, Frho, e_v, p_v):
        n = C.c_int()
        z = _lib.orc_evolve_ellipse(self.dlin.buf, self.Om, self.OL, Frho, e_v, p_v, C.byref(n))
        return z, n.value

    def dfnofa(self, a):
        return _lib.orc_dfnofa(self.dlin.buf, a)

    def table(self, nx=50, ny=20, nz=20, bounds=TABLE_BOUNDS):
        out = np.zeros(nx * ny * nz, np.float32)
        _lib.orc_make_collapse_table(self.dlin.buf, self.Om, self.OL, nx, ny, nz, *bounds, out.ctypes.data)
        return out.reshape(nz, ny, nx).transpose(2, 1, 0)   # [ix, iy, iz]
