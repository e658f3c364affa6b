"""ctypes binding of libpeakpatch_hip.so (include/peakpatch_hip.h).

The shared library is built in-tree by ``peakpatch.jl_amd/build.py`` into
``peakpatch.jl_amd/lib/libpeakpatch_hip.so``.  Loading fails loudly when it is
missing: there is no CPU fallback behind this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PEAKPATCH_HIP_LIB") or os.path.join(os.path.dirname(HERE), "lib", "libpeakpatch_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include", "peakpatch_hip.h")

PP_OK = 0
ERRORS = {-1: "PP_ERR_ARG", -2: "PP_ERR_OOM", -3: "PP_ERR_HIP", -4: "PP_ERR_FFT", -5: "PP_ERR_STATE",
          -6: "PP_ERR_CAPACITY", -7: "PP_ERR_BUSY", -8: "PP_ERR_TIMEOUT", -9: "PP_ERR_INJECTED"}
SHELL_PLAIN, SHELL_PERSISTENT = 0, 1


class PPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Peaks(C.Structure):
    _fields_ = [("capacity", C.c_int64), ("count", C.c_int64), ("needed", C.c_int64),
                ("i", C.c_void_p), ("j", C.c_void_p), ("k", C.c_void_p), ("filt", C.c_void_p),
                ("delta_s", C.c_void_p), ("lapd_s", C.c_void_p)]


class Fields(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int),
                ("delta", C.c_void_p), ("etax", C.c_void_p), ("etay", C.c_void_p), ("etaz", C.c_void_p),
                ("eta2x", C.c_void_p), ("eta2y", C.c_void_p), ("eta2z", C.c_void_p),
                ("lapd", C.c_void_p), ("on_device", C.c_int)]


class PeakBatch(C.Structure):
    _fields_ = [("npeaks", C.c_int64), ("i", C.c_void_p), ("j", C.c_void_p), ("k", C.c_void_p),
                ("Rf", C.c_void_p), ("ir2min", C.c_void_p), ("ZZon", C.c_void_p), ("fcrit", C.c_void_p),
                ("on_device", C.c_int)]


class Results(C.Structure):
    _fields_ = [("RTHL", C.c_void_p), ("Srb", C.c_void_p), ("Sbar", C.c_void_p), ("Sbar2", C.c_void_p),
                ("strain", C.c_void_p), ("gradpk", C.c_void_p), ("gradpkrf", C.c_void_p),
                ("Fbarx", C.c_void_p), ("e_v", C.c_void_p), ("p_v", C.c_void_p),
                ("zvir_half", C.c_void_p), ("gradpkf", C.c_void_p), ("d2F", C.c_void_p),
                ("status", C.c_void_p), ("nshells", C.c_void_p), ("on_device", C.c_int),
                ("dump_counts", C.c_int64 * 7)]


class HaloInputs(C.Structure):
    _fields_ = [("npeaks", C.c_int64), ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("Rf", C.c_void_p), ("ZZon", C.c_void_p), ("D", C.c_void_p), ("v2fac", C.c_void_p),
                ("FcRf", C.c_void_p), ("d2Rf", C.c_void_p), ("on_device", C.c_int)]


class Growth(C.Structure):
    _fields_ = [("ntab", C.c_int), ("atab", C.c_void_p), ("D_a", C.c_void_p), ("dlnD_dlna", C.c_void_p)]


_vp, _i, _d, _f, _i64 = C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_int64

class PkTable(C.Structure):
    _fields_ = [("npk", C.c_int), ("k", C.c_void_p), ("P", C.c_void_p)]


class TileFields(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in ("delta", "etax", "etay", "etaz", "eta2x", "eta2y", "eta2z", "lapd")]


class TileGeom(C.Structure):
    _fields_ = [("n", C.c_int), ("alatt", C.c_double), ("xbx", C.c_double), ("ybx", C.c_double),
                ("zbx", C.c_double), ("ievol", C.c_int), ("z_out", C.c_double), ("obs", C.c_double * 3),
                ("z_max", C.c_double), ("ilpt", C.c_int), ("Om", C.c_double), ("OL", C.c_double),
                ("round_f32", C.c_int), ("growth", Growth), ("nchi", C.c_int), ("chi_z", C.c_void_p),
                ("chi_chi", C.c_void_p), ("chi_zmax", C.c_double)]


class SplitGeom(C.Structure):
    _fields_ = [("nmesh", C.c_int), ("nbuff", C.c_int), ("ntile", C.c_int), ("M", C.c_int), ("alatt", C.c_double),
                ("seed", C.c_uint64)]


BATCH_FIELDS = ("i", "j", "k", "Rf", "ir2min", "ZZon", "fcrit", "x", "y", "z", "D", "v2fac", "FcRf", "d2Rf")


class BatchOut(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in BATCH_FIELDS]


# pp_tile_fn(user, ctx, slot, ticket, request) -> status
TILE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int64, C.c_void_p)


class ServeConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("slots", C.c_int), ("queue_capacity", C.c_int),
                ("budget_bytes", C.c_size_t), ("hbm_fraction", C.c_double), ("fn", TILE_FN), ("user", C.c_void_p)]


SIGNATURES = {
    "pp_last_error": (C.c_char_p, []),
    "pp_version": (_i, []),
    "pp_device_count": (_i, [C.POINTER(_i)]),
    "pp_ctx_create": (_i, [_i, C.POINTER(_vp)]),
    "pp_ctx_destroy": (_i, [_vp]),
    "pp_ctx_synchronize": (_i, [_vp]),
    "pp_ctx_mem_info": (_i, [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "pp_ctx_stream": (_i, [_vp, C.POINTER(_vp)]),
    "pp_ctx_wait_stream": (_i, [_vp, _vp]),
    "pp_ctx_signal_stream": (_i, [_vp, _vp]),
    "pp_upload_collapse_table": (_i, [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _d, _d]),
    "pp_make_collapse_table": (_i, [_vp, _d, _d, C.POINTER(Growth), _i, _i, _i, _d, _d, _d, _d, _d, _d, _vp, _i]),
    "pp_build_shell_tables": (_i, [_vp, _i]),
    "pp_shell_table_info": (_i, [_vp, C.POINTER(_i64), C.POINTER(C.c_int32)]),
    "pp_window_multiply": (_i, [_vp, _vp, _vp, _i, _i, _d, _d, _i, _f]),
    "pp_smooth_field": (_i, [_vp, _vp, _vp, _i, _i, _d, _d, _i]),
    "pp_rfft": (_i, [_vp, _vp, _vp, _i, _i]),
    "pp_find_peaks_field": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _f, _i64, C.POINTER(Peaks)]),
    "pp_find_peaks_field_noalloc": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _f, _i64, C.POINTER(Peaks)]),
    "pp_find_peaks": (_i, [_vp, _vp, _i, _i, _d, _vp, _vp, _i, _i, _i, _i, _i64, _vp, C.POINTER(Peaks)]),
    "pp_analyse_peaks": (_i, [_vp, C.POINTER(Fields), C.POINTER(PeakBatch), _d, _d, _i, _vp, _i, _i,
                              C.POINTER(Results)]),
    "pp_shell_profiles": (_i, [_vp, C.POINTER(Fields), C.POINTER(PeakBatch), _d, _d, _i, _i, _vp, _vp, _vp]),
    "pp_autotune_shell": (_i, [_vp, C.POINTER(Fields), C.POINTER(PeakBatch), _d, _d, C.POINTER(_i)]),
    "pp_last_timing": (_i, [_vp, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d)]),
    "pp_shell_tunables": (_i, [_vp, _vp, _vp]),
    "pp_last_shell_stages": (_i, [_vp, C.POINTER(_d)]),
    "pp_fill_noise_threefry": (_i, [_vp, _vp, _i64, C.c_uint64]),
    "pp_make_tile_fields": (_i, [_vp, _i, _d, C.c_uint64, C.POINTER(PkTable), C.POINTER(TileFields)]),
    "pp_ctx_release_scratch": (_i, [_vp]),
    "pp_coarse_create": (_i, [_vp, C.POINTER(SplitGeom), C.POINTER(PkTable), C.POINTER(C.c_void_p)]),
    "pp_coarse_download": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "pp_coarse_destroy": (_i, [_vp]),
    "pp_split_tile_fields": (_i, [_vp, _vp, _i, _i, _i, C.POINTER(TileFields)]),
    "pp_merge_halos": (_i, [_vp, _i64, _vp, _vp, _vp, _vp, _i, _vp, _vp, C.POINTER(C.c_int64)]),
    "pp_bookkeep_peaks": (_i, [_vp, C.POINTER(Peaks), _vp, _i, C.POINTER(TileGeom), C.POINTER(BatchOut),
                               C.POINTER(C.c_int64)]),
    "pp_serve_create": (_i, [C.POINTER(ServeConfig), C.POINTER(_vp)]),
    "pp_serve_submit": (_i, [_vp, _vp, C.c_size_t, C.POINTER(_i64)]),
    "pp_serve_poll": (_i, [_vp, _i64, _i, C.POINTER(_i), C.POINTER(_d)]),
    "pp_serve_next": (_i, [_vp, _i, C.POINTER(_i64), C.POINTER(_vp), C.POINTER(_i), C.POINTER(_d)]),
    "pp_serve_message": (_i, [_vp, _i64, C.c_char_p, C.c_size_t]),
    "pp_serve_inject_fault": (_i, [_vp, _i64]),
    "pp_serve_stats": (_i, [_vp, C.POINTER(_i64)]),
    "pp_serve_context": (_i, [_vp, _i, C.POINTER(_vp)]),
    "pp_serve_destroy": (_i, [_vp]),
    "pp_pack_halos": (_i, [_vp, C.POINTER(Results), C.POINTER(HaloInputs), _d, _d, _i, _vp, _i64, _i,
                           C.POINTER(_i64), C.POINTER(_f)]),
}

_lib = None


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built; run `python peakpatch.jl_amd/build.py`")
    # PyTorch-ROCm ships its own HIP runtime.  Whichever copy loads first
    # serves the process (same SONAME), and torch cannot start its GPU on a
    # runtime loaded before it, so torch, when present, is loaded first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int):
    if rc != PP_OK:
        msg = load().pp_last_error()
        raise PPError(rc, msg.decode() if msg else "")
    return rc
