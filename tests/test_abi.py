"""The C-ABI library loads and exports every entry point its header declares
(no GPU call is made)."""
import ctypes
import re

from peakpatch_hip import _lib as L


def declared_functions():
    src = open(L.HEADER_PATH).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(pp_[a-z_0-9]+)\s*\(", src, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("pp_ctx_create", "pp_find_peaks", "pp_analyse_peaks", "pp_smooth_field",
                 "pp_upload_collapse_table", "pp_build_shell_tables", "pp_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_the_header():
    assert set(declared_functions()) <= set(L.SIGNATURES)
    L.load()


def test_last_error_is_empty_string_initially():
    assert L.load().pp_last_error() is not None
