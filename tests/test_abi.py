"""The C ABI: libimdg.so loads without a GPU and exports every entry point
that include/imdg.h declares; the ctypes structures match the header's
layout; errors come back as codes, not crashes, for calls that need no
device (no compute call is made here)."""
import ctypes as C
import os
import re

import pytest

from imd_amd import build as B
from imd_amd import imdg as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "imdg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(imdg_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(B.LIB):
        B.build_lib()
    return C.CDLL(B.LIB)


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("imdg_calc_forces", "imdg_move_atoms", "imdg_make_nblist",
              "imdg_check_nblist", "imdg_deallocate_nblist", "imdg_run",
              "imdg_set_tables", "imdg_set_atoms"):
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_library(lib):
    # the library reports sizeof() of each public struct
    lib.imdg_struct_size.argtypes = [C.c_char_p]
    lib.imdg_struct_size.restype = C.c_long
    for name, cls in (("imdg_config", G.Config), ("imdg_table", G.Table),
                      ("imdg_replica", G.Replica), ("imdg_thermo", G.Thermo),
                      ("imdg_info", G.Info), ("imdg_tuning", G.Tuning)):
        assert lib.imdg_struct_size(name.encode()) == C.sizeof(cls), name
    assert lib.imdg_struct_size(b"nonsense") == -1


def test_null_arguments_are_rejected(lib):
    lib.imdg_ctx_create.argtypes = [C.c_void_p, C.c_void_p]
    lib.imdg_calc_forces.argtypes = [C.c_void_p, C.c_int]
    lib.imdg_last_error.restype = C.c_char_p
    lib.imdg_last_error.argtypes = [C.c_void_p]
    assert lib.imdg_ctx_create(None, None) == G.ERR_ARG
    assert lib.imdg_calc_forces(None, 0) == G.ERR_ARG
    assert lib.imdg_last_error(None)
