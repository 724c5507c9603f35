"""The C ABI library loads and exports every entry point the header declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

from pfem2_amd import EXPORTS, LIB_PATH, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "pfem2_amd.h")).read()
    return sorted(set(re.findall(r"\b(pf2_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert header_functions() == sorted(EXPORTS)


def test_library_exports_all_symbols():
    assert os.path.exists(LIB_PATH), "build the library first (python -m pfem2_amd.build)"
    lib = load_library()
    for name in header_functions():
        assert hasattr(lib, name), name


def test_error_string_without_gpu_work():
    lib = load_library()
    h = ctypes.c_void_p()
    # NULL mesh is rejected before any device call
    rc = lib.pf2_create(None, None, 0, ctypes.byref(h))
    assert rc == -1
    assert b"NULL" in lib.pf2_last_error()
    assert lib.pf2_kernel_name(0) == b"k_move"
