"""CPU-side checks of the C ABI: the library loads and exports every
function include/gbs_hip.h declares; argument errors come back as status
codes with a message (no GPU needed for these)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(gbs_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "gbs_gemm" in names and "gbs_import" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    import gemmbitserial_amd as g
    L = g.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(g.EXPORTED) == declared_functions()


def test_version_and_errors_without_device():
    import gemmbitserial_amd as g
    L = g.lib()
    assert L.gbs_version() >= 10000
    out = C.c_void_p()
    # argument validation happens before any device work
    assert L.gbs_matrix_create(9, 4, 4, 0, 4, 64, C.byref(out)) == 1
    assert b"nbits" in L.gbs_last_error()
    assert L.gbs_matrix_create(2, 4, 70, 0, 4, 100, C.byref(out)) == 1
    assert b"multiple of 64" in L.gbs_last_error()
    assert L.gbs_matrix_create(2, 4, 4, 0, 4, 64, None) == 1
    assert L.gbs_gemm(None, None, None, None, 0, 0) == 1
    assert b"null matrix" in L.gbs_last_error()


def test_import_refuses_unknown_dtype():
    import gemmbitserial_amd as g
    L = g.lib()
    # a fake non-null matrix pointer is never dereferenced before the dtype check
    assert L.gbs_import(C.c_void_p(8), C.c_void_p(16), 42, 0, 0, None) == 1
    assert b"dtype" in L.gbs_last_error()
