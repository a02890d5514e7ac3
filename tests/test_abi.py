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


def test_conv2d_shape_closed_form():
    import gemmbitserial_amd as g
    for (N, Cc, H, W, kh, kw, s, p, d) in [(16, 256, 56, 56, 3, 3, 1, 1, 1),
                                         (1, 3, 227, 227, 11, 11, 4, 0, 1),
                                         (2, 5, 9, 7, 3, 2, (2, 1), (1, 0), (1, 2)),
                                         (1, 1, 1, 1, 1, 1, 1, 0, 1)]:
        geo = g.conv2d(N, Cc, H, W, kh, kw, s, p, d)
        sh, sw = (s, s) if isinstance(s, int) else s
        ph, pw = (p, p) if isinstance(p, int) else p
        dh, dw = (d, d) if isinstance(d, int) else d
        oh = (H + 2 * ph - (dh * (kh - 1) + 1)) // sh + 1
        ow = (W + 2 * pw - (dw * (kw - 1) + 1)) // sw + 1
        assert g.conv2d_shape(geo) == (N * oh * ow, Cc * kh * kw, oh, ow)
    # the CaffeNet conv1 operand: 55 x 55 output pixels, K = 3 * 11 * 11 = 363
    assert g.conv2d_shape(g.conv2d(1, 3, 227, 227, 11, 11, 4))[:2] == (3025, 363)
    for bad in [g.conv2d(1, 1, 2, 2, 3, 3), g.conv2d(0, 1, 4, 4, 1, 1),
                g.conv2d(1, 1, 4, 4, 1, 1, stride=0)]:
        with pytest.raises(g.GbsError):
            g.conv2d_shape(bad)


def test_matrix_load_rejects_bad_files_before_touching_the_device(tmp_path):
    import numpy as np
    import gemmbitserial_amd as g
    L = g.lib()
    out = C.c_void_p()
    good_head = np.array([2, 5, 70, 8, 128, 1, 2 * 8 * 2], dtype="<u8")
    cases = {
        "magic": b"NOTPLANE" + good_head.tobytes() + bytes(8 * 32),
        "short": g.PLANE_FILE_MAGIC + good_head.tobytes() + bytes(8 * 31),
        "count": g.PLANE_FILE_MAGIC + np.array([2, 5, 70, 8, 128, 1, 31], "<u8").tobytes()
        + bytes(8 * 31),
        "ncols_a": g.PLANE_FILE_MAGIC + np.array([2, 5, 70, 8, 100, 1, 24], "<u8").tobytes()
        + bytes(8 * 24),
        "nbits": g.PLANE_FILE_MAGIC + np.array([9, 5, 70, 8, 128, 1, 144], "<u8").tobytes()
        + bytes(8 * 144),
    }
    for name, blob in cases.items():
        p = tmp_path / f"{name}.gbsp"
        p.write_bytes(blob)
        assert L.gbs_matrix_load(os.fsencode(str(p)), C.byref(out)) == 1, name
        msg = L.gbs_last_error()
        assert b"not a packed-matrix file" in msg or b"inconsistent" in msg, (name, msg)
    assert L.gbs_matrix_load(b"/nonexistent/x.gbsp", C.byref(out)) == 1


def test_batch_planes_mode_without_device():
    """Which products batches put into the planes group: copies by default,
    modes 0-2 set and read back, anything else refused with the mode named."""
    import gemmbitserial_amd as g
    L = g.lib()
    assert g.batch_get_planes() == g.PLANES_NONE
    try:
        for m in (g.PLANES_SKINNY, g.PLANES_ALL, g.PLANES_NONE):
            g.batch_set_planes(m)
            assert g.batch_get_planes() == m
        assert L.gbs_batch_set_planes(3) == 1 and b"mode 3" in L.gbs_last_error()
        assert L.gbs_batch_set_planes(-1) == 1
        assert g.batch_get_planes() == g.PLANES_NONE
        assert L.gbs_batch_get_planes(None) == 1
    finally:
        g.batch_set_planes(g.PLANES_NONE)


def test_batch_tune_table_validation_without_device():
    """The persistent launch's tile table: entries are checked against the
    shapes the persistent kernel carries per format, and the committed table
    (written by the tuner) loads."""
    import gemmbitserial_amd as g
    n = g.batch_tune_classes()
    assert n == 18
    assert g.batch_tune_class_of(1, 256) == 0 and g.batch_tune_class_of(256, 8192) == 14
    assert g.batch_tune_class_of(257, 1024) == 16
    for fmt in range(3):
        shapes = g.batch_tune_shapes(fmt)
        assert shapes and 13 in shapes or 18 in shapes
    assert g.batch_tune_shapes(2) == [2, 18]
    L = g.lib()
    saved = [[g.batch_tune_get(f, c) for c in range(n)] for f in range(3)]
    try:
        assert L.gbs_batch_tune_set(2, 0, 13, 16) == 1      # 64x64 dl4 is not an FP6 shape
        assert b"not carried" in L.gbs_last_error()
        assert L.gbs_batch_tune_set(1, 0, 13, 0) == 1       # piece out of range
        assert L.gbs_batch_tune_set(1, n, 13, 16) == 1      # class out of range
        g.batch_tune_set(1, 4, 2, 8)
        assert g.batch_tune_get(1, 4) == (2, 8)
        g.batch_tune_set(1, 4, -1)
        assert g.batch_tune_get(1, 4) == (-1, 0)
        table = os.path.join(ROOT, "gemmbitserial_amd", "batch_tune_table.txt")
        if os.path.exists(table):
            g.batch_tune_load(table)
            got = [g.batch_tune_get(f, c) for f in range(3) for c in range(n)]
            assert any(s >= 0 for s, _ in got)
    finally:
        for f in range(3):
            for c in range(n):
                s, p = saved[f][c]
                g.batch_tune_set(f, c, s, p if s >= 0 else 16)


def test_batch_tuner_table_round_trip(tmp_path):
    """The table the batch tuner writes is the one the library loads: every
    entry comes back, classes not written revert to the built-in rule."""
    import gemmbitserial_amd as g
    from gemmbitserial_amd import tune
    n = g.batch_tune_classes()
    saved = [[g.batch_tune_get(f, c) for c in range(n)] for f in range(3)]
    entries = {(1, 0): (3, 16), (2, 7): (18, 32), (1, 14): (13, 8), (0, 4): (10, 16)}
    detail = {"workload": "test", "builtin_us": 1.0, "tuned_us": 0.9,
              "products": {k: 1 for k in entries}, "mbytes": {k: 0.1 for k in entries}}
    path = tmp_path / "batch_tune_table.txt"
    tune.write_batch_table(entries, detail, str(path))
    try:
        g.batch_tune_load(str(path))
        for f in range(3):
            for c in range(n):
                assert g.batch_tune_get(f, c) == entries.get((f, c), (-1, 0)), (f, c)
        bad = tmp_path / "bad.txt"
        bad.write_text("0 2 13 16\n")        # 64x64 dl4 is not an FP6 shape
        L = g.lib()
        assert L.gbs_batch_tune_load(str(bad).encode()) == 1
        assert b"invalid" in L.gbs_last_error()
    finally:
        for f in range(3):
            for c in range(n):
                s, p = saved[f][c]
                g.batch_tune_set(f, c, s, p if s >= 0 else 16)


def test_serving_tuner_rewrites_only_int8_entries(tmp_path):
    """tune --serving rewrites the int8 lines of the batch table and keeps
    every other format's line; the result loads."""
    import gemmbitserial_amd as g
    from gemmbitserial_amd import tune
    path = tmp_path / "t.txt"
    path.write_text("# header\n3 1 3 16  # fp4\n4 2 18 16  # fp6\n1 0 2 16  # old int8\n"
                    "# serving (int8) entries, old\n2 0 3 8  # old\n")
    tune.write_serving_entries({1: (13, 4), 2: (13, 2)}, {"workload": "w", "tuned_ms": 0.1,
                                                          "builtin_ms": 0.2}, str(path))
    lines = [ln.split("#")[0].split() for ln in path.read_text().splitlines()]
    rows = [tuple(int(x) for x in ln) for ln in lines if ln]
    assert sorted(rows) == sorted([(3, 1, 3, 16), (4, 2, 18, 16), (1, 0, 13, 4), (2, 0, 13, 2)])
    n = g.batch_tune_classes()
    saved = [[g.batch_tune_get(f, c) for c in range(n)] for f in range(3)]
    try:
        g.batch_tune_load(str(path))
        assert g.batch_tune_get(0, 1) == (13, 4) and g.batch_tune_get(1, 3) == (3, 16)
    finally:
        for f in range(3):
            for c in range(n):
                s, p = saved[f][c]
                g.batch_tune_set(f, c, s, p if s >= 0 else 16)
