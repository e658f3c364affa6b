"""Host-side front end of the C-ABI: one ``Context`` per GPU.

Arrays follow the reference's memory layout: a Julia ``Array{Float32,3}`` of
size (n, n, n) is x-fastest, so the numpy/torch view is C-order with shape
(n, n, n) indexed [k, j, i].  Grid indices returned here are 0-based.

Inputs may be numpy arrays (host) or torch CUDA tensors (device memory of the
context's GPU); outputs of the peak and shell entry points are numpy arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

RESULT_FIELDS = ("RTHL", "Srb", "Sbar", "Sbar2", "strain", "gradpk", "gradpkrf", "Fbarx", "e_v", "p_v",
                 "zvir_half", "gradpkf", "d2F")
_VEC3 = {"Sbar", "Sbar2", "gradpk", "gradpkrf", "gradpkf"}
DUMP_NAMES = ("none", "premask", "no_fcrit", "m0_one", "no_collapse", "rthl_neg", "nonfinite")


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _ptr(a, dtype, what: str):
    """Return (address, on_device, keepalive) for a numpy array or torch tensor."""
    if a is None:
        return None, 0, None
    if _is_torch(a):
        import torch
        want = {np.float32: torch.float32, np.int8: torch.int8, np.complex64: torch.complex64}[dtype]
        if a.dtype != want or not a.is_contiguous():
            raise ValueError(f"{what}: torch tensor must be contiguous {want}")
        return a.data_ptr(), 1 if a.is_cuda else 0, a
    arr = np.ascontiguousarray(a, dtype=dtype)
    return arr.ctypes.data, 0, arr


@dataclass
class PeakList:
    """Output of the peak finder: 0-based indices plus per-peak filter data."""
    i: np.ndarray
    j: np.ndarray
    k: np.ndarray
    filt: np.ndarray
    delta_s: np.ndarray
    lapd_s: np.ndarray | None

    def __len__(self):
        return len(self.i)


class Context:
    """A pp_ctx: device, stream, rocFFT plans, shell/collapse tables."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        L.check(self.lib.pp_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device
        self.rmax = None
        self._owned = True

    @classmethod
    def from_handle(cls, handle: int, device: int = 0) -> "Context":
        """A non-owning wrapper of an existing pp_ctx (e.g. a serving slot's
        context); close() leaves the context to its owner."""
        self = cls.__new__(cls)
        self.lib = L.load()
        self.h = C.c_void_p(handle)
        self.device = device
        self.rmax = None
        self._owned = False
        return self

    def close(self):
        if self.h:
            if self._owned:
                self.lib.pp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ tables
    def upload_collapse_table(self, table: np.ndarray, X1, X2, Y1, Y2, Z1, Z2):
        """table: (Nx, Ny, Nz) array indexed [ix, iy, iz] (x fastest in memory)."""
        t = np.asarray(table, np.float32)
        nx, ny, nz = t.shape
        flat = np.ascontiguousarray(t.ravel(order="F"))
        L.check(self.lib.pp_upload_collapse_table(self.h, flat.ctypes.data, nx, ny, nz,
                                                  X1, X2, Y1, Y2, Z1, Z2))

    def build_shell_tables(self, rmax: int):
        L.check(self.lib.pp_build_shell_tables(self.h, int(rmax)))
        self.rmax = int(rmax)

    def shell_table_info(self):
        nc = C.c_int64(); ns = C.c_int32()
        L.check(self.lib.pp_shell_table_info(self.h, C.byref(nc), C.byref(ns)))
        return nc.value, ns.value

    def mem_info(self):
        f = C.c_size_t(); t = C.c_size_t()
        L.check(self.lib.pp_ctx_mem_info(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def stream_handle(self) -> int:
        """The context's HIP stream as an integer handle
        (torch.cuda.ExternalStream accepts it)."""
        s = C.c_void_p()
        L.check(self.lib.pp_ctx_stream(self.h, C.byref(s)))
        return int(s.value or 0)

    def _call(self, fn, on_device):
        """Run one library call.  With device operands the context's streams
        first wait for the work queued so far on torch's current stream
        (where the caller's tensors were produced), and afterwards that
        stream waits for the context's (where they were written), so torch
        work on either side of the call is ordered with it."""
        if not on_device:
            return L.check(fn())
        import torch
        s = torch.cuda.current_stream(self.device).cuda_stream
        L.check(self.lib.pp_ctx_wait_stream(self.h, s))
        rc = L.check(fn())
        L.check(self.lib.pp_ctx_signal_stream(self.h, s))
        return rc

    def release_scratch(self):
        """Return the context's pooled working memory to the device (tables
        kept); waits for the context's streams."""
        L.check(self.lib.pp_ctx_release_scratch(self.h))

    def fill_noise(self, n: int, seed: int, out=None):
        """Threefry white noise of an n^3 tile as a CUDA float32 tensor
        (RandomField.jl:46-81), generated on the context's stream and
        ordered before later work on torch's current stream."""
        import torch
        if out is None:
            out = torch.empty((n, n, n), dtype=torch.float32, device=f"cuda:{self.device}")
        self._call(lambda: self.lib.pp_fill_noise_threefry(self.h, out.data_ptr(), n * n * n,
                                                           int(seed) & 0xFFFFFFFFFFFFFFFF), True)
        return out

    def synchronize(self):
        L.check(self.lib.pp_ctx_synchronize(self.h))

    TUNABLES = ("small_m", "chunk_shells", "coop_m", "gather_waves_cu", "post_width", "lane_walk")

    def tunables(self, **kw):
        """Get (and with keywords, set) the persistent form's launch
        parameters: small_m, chunk_shells, coop_m, gather_waves_cu, post_width,
        lane_walk."""
        unknown = set(kw) - set(self.TUNABLES)
        if unknown:
            raise ValueError(f"unknown tunables {sorted(unknown)}")
        get = np.zeros(len(self.TUNABLES), np.int32)
        if kw:
            st = np.array([-1, 0, 0, 0, 0, -1], np.int32)
            for q, name in enumerate(self.TUNABLES):
                if name in kw:
                    st[q] = int(kw[name])
            L.check(self.lib.pp_shell_tunables(self.h, st.ctypes.data, get.ctypes.data))
        else:
            L.check(self.lib.pp_shell_tunables(self.h, None, get.ctypes.data))
        return dict(zip(self.TUNABLES, (int(v) for v in get)))

    STAGES = ("prepare", "gather", "small_gather", "post_wave", "post_lane", "finish", "small_post")

    def last_timing(self):
        g = C.c_double(); s = C.c_double(); p = C.c_double()
        L.check(self.lib.pp_last_timing(self.h, C.byref(g), C.byref(s), C.byref(p)))
        st = (C.c_double * len(self.STAGES))()
        L.check(self.lib.pp_last_shell_stages(self.h, st))
        return {"gather_ms": g.value, "shell_ms": s.value, "peakfind_ms": p.value,
                "stages_ms": dict(zip(self.STAGES, list(st)))}

    # ------------------------------------------------------- filter bank
    def window_multiply(self, delta_k, n: int, boxsize: float, Rf: float, wsmooth: int, scale: float = 1.0):
        """k-space multiply of smooth_field on a host complex64 half-spectrum
        shaped (n, n, n//2+1); returns a new numpy array."""
        a = np.ascontiguousarray(delta_k, np.complex64)
        out = np.empty_like(a)
        L.check(self.lib.pp_window_multiply(self.h, a.ctypes.data, out.ctypes.data, 0, n, boxsize, Rf,
                                            wsmooth, scale))
        return out

    def rfft(self, field) -> np.ndarray:
        f = np.ascontiguousarray(field, np.float32)
        n = f.shape[0]
        out = np.empty((n, n, n // 2 + 1), np.complex64)
        L.check(self.lib.pp_rfft(self.h, f.ctypes.data, out.ctypes.data, 0, n))
        return out

    def smooth_field(self, delta_k, n: int, boxsize: float, Rf: float, wsmooth: int) -> np.ndarray:
        """smooth_field (Filters.jl:79-106) on the GPU; host in, host out."""
        a = np.ascontiguousarray(delta_k, np.complex64)
        out = np.empty((n, n, n), np.float32)
        L.check(self.lib.pp_smooth_field(self.h, a.ctypes.data, out.ctypes.data, 0, n, boxsize, Rf, wsmooth))
        return out

    def find_peaks_field(self, delta_s, mask, nbuff: int, fcrit: float, max_peaks: int = 2_000_000,
                         capacity: int | None = None) -> PeakList:
        """find_peaks (PeakFind.jl:15-78) on one field; mask (int8, same
        shape) is updated in place.  Host numpy or device torch inputs."""
        dp, dev, keep1 = _ptr(delta_s, np.float32, "delta_s")
        if not _is_torch(mask):
            if not (isinstance(mask, np.ndarray) and mask.dtype == np.int8 and mask.flags.c_contiguous):
                raise ValueError("mask must be a C-contiguous int8 numpy array (updated in place)")
            mp = mask.ctypes.data
        else:
            mp, _, _ = _ptr(mask, np.int8, "mask")
        n3, n2, n1 = delta_s.shape
        cap = capacity if capacity is not None else max(1, min(max_peaks, (n1 - 2 * nbuff) * (n2 - 2 * nbuff) * (n3 - 2 * nbuff)))
        return self._run_peaks(lambda pk: self.lib.pp_find_peaks_field(
            self.h, dp, mp, dev, n1, n2, n3, nbuff, fcrit, max_peaks, C.byref(pk)), cap, False, dev)

    def find_peaks(self, delta, boxsize: float, Rf, fcrit, wsmooth: int = 1, want_lapd: bool = False,
                   nbuff: int = 1, max_peaks_per_filter: int = 2_000_000, mask=None,
                   capacity: int | None = None, device_out: bool = False) -> PeakList:
        """Filter bank + stencil over all filters in the given order
        (peak_find_tile_gpu, ext/CUDAExt.jl:3068-3197).  device_out: the
        peak arrays are CUDA tensors (for bookkeep_peaks) instead of numpy."""
        dp, dev, keep = _ptr(delta, np.float32, "delta")
        n = delta.shape[0]
        Rf = np.ascontiguousarray(Rf, np.float64)
        fc = np.ascontiguousarray(np.broadcast_to(np.asarray(fcrit, np.float32), Rf.shape), np.float32)
        mp = None
        if mask is not None:
            if not (isinstance(mask, np.ndarray) and mask.dtype == np.int8 and mask.flags.c_contiguous):
                raise ValueError("mask must be a C-contiguous int8 numpy array (updated in place)")
            mp = mask.ctypes.data
        inner = max(0, n - 2 * nbuff) ** 3
        cap = capacity if capacity is not None else max(1, min(inner, len(Rf) * max_peaks_per_filter, 1 << 24))
        return self._run_peaks(lambda pk: self.lib.pp_find_peaks(
            self.h, dp, dev, n, boxsize, Rf.ctypes.data, fc.ctypes.data, len(Rf), wsmooth, int(want_lapd),
            nbuff, max_peaks_per_filter, mp, C.byref(pk)), cap, want_lapd, dev, device_out)

    def _run_peaks(self, call, cap, want_lapd, on_device, device_out=False) -> PeakList:
        names = ("i", "j", "k", "filt", "delta_s", "lapd_s")
        if device_out:
            import torch
            d = torch.device("cuda", self.device)
            bufs = {k: torch.empty(cap, dtype=torch.int32, device=d) for k in names[:4]}
            bufs["delta_s"] = torch.empty(cap, dtype=torch.float32, device=d)
            bufs["lapd_s"] = torch.empty(cap, dtype=torch.float32, device=d) if want_lapd else None
            addr = lambda b: b.data_ptr()
        else:
            bufs = {k: np.zeros(cap, np.int32) for k in names[:4]}
            bufs["delta_s"] = np.zeros(cap, np.float32)
            bufs["lapd_s"] = np.zeros(cap, np.float32) if want_lapd else None
            addr = lambda b: b.ctypes.data
        pk = L.Peaks(cap, 0, 0, *(addr(bufs[k]) if bufs[k] is not None else None for k in names))
        self._call(lambda: call(pk), on_device or device_out)
        c = pk.count
        cut = (lambda b: b[:c]) if device_out else (lambda b: b[:c].copy())
        return PeakList(*(cut(bufs[k]) if bufs[k] is not None else None for k in names))

    def bookkeep_peaks(self, peaks: PeakList, Rf_filters, geom: L.TileGeom) -> dict:
        """The drivers' per-peak loop (src/Pipeline.jl:217-231, 257-271;
        src/MultiResolution.jl:959-988) on the device, over a device PeakList
        of find_peaks(device_out=True).  Returns the kept rows as CUDA tensors
        keyed like L.BATCH_FIELDS (Rf, ir2min, ZZon, fcrit, x, y, z, D, v2fac
        Float64; FcRf, d2Rf Float32; i, j, k int32)."""
        import torch
        n = len(peaks)
        d = torch.device("cuda", self.device)
        dt = {"i": torch.int32, "j": torch.int32, "k": torch.int32, "ir2min": torch.int64,
              "FcRf": torch.float32, "d2Rf": torch.float32}
        out = {f: torch.empty(max(n, 1), dtype=dt.get(f, torch.float64), device=d) for f in L.BATCH_FIELDS}
        if n == 0:
            return {f: t[:0] for f, t in out.items()}
        for a in (peaks.i, peaks.j, peaks.k, peaks.filt, peaks.delta_s):
            if not (_is_torch(a) and a.is_cuda):
                raise ValueError("bookkeep_peaks needs the device PeakList of find_peaks(device_out=True)")
        pk = L.Peaks(n, n, n, *(a.data_ptr() if a is not None else None
                                for a in (peaks.i, peaks.j, peaks.k, peaks.filt, peaks.delta_s, peaks.lapd_s)))
        rf = np.ascontiguousarray(Rf_filters, np.float64)
        bo = L.BatchOut(*(out[f].data_ptr() for f in L.BATCH_FIELDS))
        kept = C.c_int64(0)
        self._call(lambda: self.lib.pp_bookkeep_peaks(self.h, C.byref(pk), rf.ctypes.data, len(rf), C.byref(geom),
                                                       C.byref(bo), C.byref(kept)), True)
        return {f: t[:kept.value] for f, t in out.items()}

    # ----------------------------------------------------- shell analysis
    def _fields(self, delta, eta, eta2=None, lapd=None):
        keep = []
        ptrs = []
        on_dev = None
        arrs = [delta, *eta, *(eta2 if eta2 is not None else (None, None, None)), lapd]
        for a in arrs:
            p, dev, k = _ptr(a, np.float32, "field")
            if a is not None:
                on_dev = dev if on_dev is None else on_dev
                if dev != on_dev:
                    raise ValueError("fields must be all host or all device")
            ptrs.append(p)
            keep.append(k)
        n3, n2, n1 = delta.shape
        return L.Fields(n1, n2, n3, *ptrs, on_dev or 0), keep

    @staticmethod
    def _batch(i, j, k, Rf, ir2min, ZZon, fcrit):
        """Host numpy batch, or a device batch when i is a CUDA tensor (then every
        per-peak array must be a CUDA tensor of the right dtype)."""
        if _is_torch(i) and i.is_cuda:
            import torch
            n = i.numel()
            want = (torch.int32, torch.int32, torch.int32, torch.float64, torch.int64, torch.float64, torch.float64)
            arrs = [i, j, k, Rf, ir2min, ZZon, fcrit]
            for a, w in zip(arrs, want):
                if not (_is_torch(a) and a.is_cuda and a.dtype == w and a.is_contiguous() and a.numel() == n):
                    raise ValueError("device batch: i, j, k int32; Rf float64; ir2min int64; ZZon, fcrit float64")
            return L.PeakBatch(n, *(a.data_ptr() for a in arrs), 1), arrs
        n = len(i)
        arrs = [np.ascontiguousarray(i, np.int32), np.ascontiguousarray(j, np.int32),
                np.ascontiguousarray(k, np.int32),
                np.ascontiguousarray(np.broadcast_to(Rf, (n,)), np.float64),
                np.ascontiguousarray(np.broadcast_to(ir2min, (n,)), np.int64),
                np.ascontiguousarray(np.broadcast_to(ZZon, (n,)), np.float64),
                np.ascontiguousarray(np.broadcast_to(fcrit, (n,)), np.float64)]
        return L.PeakBatch(n, *(a.ctypes.data for a in arrs), 0), arrs

    def analyse_peaks(self, delta, eta, i, j, k, Rf, ir2min, ZZon, fcrit, alatt: float, eta2=None,
                      lapd=None, rmax2rs: float = 0.0, nbuff: int = 0, mask=None,
                      kernel: int = L.SHELL_PLAIN, device_out: bool = False) -> dict:
        """analyse_peak over a batch (RadialShell.jl:282-891).  Returns a dict
        of arrays keyed as PeakResult plus 'status', 'nshells' and
        'dump_counts'; numpy on the host, or CUDA tensors with device_out.
        mask (int8 numpy or device tensor) is updated."""
        f, keep = self._fields(delta, eta, eta2, lapd)
        b, keepb = self._batch(i, j, k, Rf, ir2min, ZZon, fcrit)
        n = int(b.npeaks)
        shapes = {name: (n, 3) if name in _VEC3 else ((n, 9) if name == "strain" else (n,)) for name in RESULT_FIELDS}
        if device_out:
            import torch
            out = {name: torch.empty(shp, dtype=torch.float64, device="cuda") for name, shp in shapes.items()}
            out["status"] = torch.empty(n, dtype=torch.int32, device="cuda")
            out["nshells"] = torch.empty(n, dtype=torch.int32, device="cuda")
            addr = lambda a: a.data_ptr()
        else:
            out = {name: np.zeros(shp, np.float64) for name, shp in shapes.items()}
            out["status"] = np.zeros(n, np.int32)
            out["nshells"] = np.zeros(n, np.int32)
            addr = lambda a: a.ctypes.data
        r = L.Results(*(addr(out[name]) for name in RESULT_FIELDS), addr(out["status"]), addr(out["nshells"]),
                      1 if device_out else 0)
        mp, mdev = None, 0
        if mask is not None:
            if _is_torch(mask):
                mp, mdev, _ = _ptr(mask, np.int8, "mask")
            else:
                if not (isinstance(mask, np.ndarray) and mask.dtype == np.int8 and mask.flags.c_contiguous):
                    raise ValueError("mask must be a C-contiguous int8 numpy array (updated in place)")
                mp = mask.ctypes.data
        self._call(lambda: self.lib.pp_analyse_peaks(self.h, C.byref(f), C.byref(b), alatt, rmax2rs, nbuff, mp,
                                                     mdev, kernel, C.byref(r)),
                   bool(f.on_device or b.on_device or mdev or device_out))
        # strain is column-major per peak: [p, L + 3K] -> (p, L, K)
        if device_out:
            out["strain"] = out["strain"].reshape(n, 3, 3).transpose(1, 2)
        else:
            out["strain"] = out["strain"].reshape(n, 3, 3, order="C").transpose(0, 2, 1)
        out["dump_counts"] = dict(zip(DUMP_NAMES, list(r.dump_counts)))
        return out

    def autotune_shell(self, delta, eta, i, j, k, Rf, ir2min, ZZon, fcrit, alatt: float, eta2=None, lapd=None,
                       rmax2rs: float = 0.0) -> dict:
        """Time the persistent form on this batch over its launch parameters
        and keep the fastest in the context; returns the chosen tunables."""
        f, keep = self._fields(delta, eta, eta2, lapd)
        b, keepb = self._batch(i, j, k, Rf, ir2min, ZZon, fcrit)
        choice = C.c_int()
        self._call(lambda: self.lib.pp_autotune_shell(self.h, C.byref(f), C.byref(b), alatt, rmax2rs,
                                                      C.byref(choice)), bool(f.on_device or b.on_device))
        return self.tunables()

    def shell_profiles(self, delta, eta, i, j, k, Rf, alatt: float, eta2=None, lapd=None,
                       rmax2rs: float = 0.0, max_shells: int = 64, kernel: int = L.SHELL_PLAIN):
        f, keep = self._fields(delta, eta, eta2, lapd)
        b, keepb = self._batch(i, j, k, Rf, 0, 1.0, 1.0)
        n = len(keepb[0])
        m = np.zeros(n, np.int32)
        ns = np.zeros((n, max_shells), np.int32)
        fb = np.zeros((n, max_shells), np.float64)
        self._call(lambda: self.lib.pp_shell_profiles(self.h, C.byref(f), C.byref(b), alatt, rmax2rs, kernel,
                                                      max_shells, m.ctypes.data, ns.ctypes.data, fb.ctypes.data),
                   bool(f.on_device))
        return m, ns, fb
