"""The C-ABI library loads without a GPU and exports every symbol that
include/omega_h_amd.h declares; the error path reports through
osh_last_error without aborting."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "omega_h_amd.h")
LIB = os.path.join(ROOT, "omega_h_amd", "libomega_h_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(osh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ("osh_adapt", "osh_build_box", "osh_measure_edges_metric",
                     "osh_measure_qualities", "osh_refine_by_size", "osh_coarsen_by_size",
                     "osh_swap_edges", "osh_coarsen_slivers", "osh_last_error"):
        assert required in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_errors_are_reported_not_aborted():
    lib = C.CDLL(LIB)
    lib.osh_last_error.restype = C.c_char_p
    lib.osh_default_opts.argtypes = [C.c_int, C.c_void_p]
    buf = C.create_string_buffer(128)
    assert lib.osh_default_opts(5, buf) != 0
    assert b"dimension" in lib.osh_last_error()
    lib.osh_mesh_nents.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    out = C.c_int()
    assert lib.osh_mesh_nents(None, 0, C.byref(out)) != 0
    assert b"null" in lib.osh_last_error()


def test_package_refuses_to_run_without_library(tmp_path):
    import omega_h_amd as A
    with pytest.raises(A.OshError):
        A.load(str(tmp_path / "missing.so"))
