"""The C++ drop-in header (include/gemmbitserial.hpp) against the reference's
own test programme shape (tests/cpp/test_dropin.cpp)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, mnist

BIN = os.path.join(ROOT, "gemmbitserial_amd", "bin", "test_dropin")


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "gemmbitserial.hpp"\n'
                   "using namespace gemmbitserial;\n"
                   "int main(){ GEMMContext c = allocateGEMMContext(2,3,4,1,1,true,true);"
                   " c.lhs.importRegular((const signed char*)0); gemmBitSerial(c);"
                   " sumRows(c.lhs, c.res); deallocGEMMContext(c); return 0; }\n")
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-Werror",
                        "-Wno-unused-function", f"-I{ROOT}/include", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_dropin_suite(tmp_path):
    assert os.path.exists(BIN), "build the C++ tests first (python -m gemmbitserial_amd.build)"
    w, x, golden = mnist()
    path = tmp_path / "mnist.bin"
    with open(path, "wb") as f:
        f.write(np.array(w.shape, np.int32).tobytes())
        f.write(w.astype(np.int8).tobytes())
        f.write(x.astype(np.uint8).tobytes())
        f.write(golden.astype(np.int32).tobytes())
    r = subprocess.run([BIN, str(path)], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "All tests completed successfully" in r.stdout
    assert "MNIST matrix-vector using bipolar times regular: OK" in r.stdout


@pytest.mark.gpu
def test_dropin_allocated_geometry_matches_reference():
    """allocGEMMContext through the header on the device: lhsBlock, rhsBlock
    and both sides' nrows_a / ncols_a equal the compiled reference's for the
    benchmark configurations' shapes, MNIST and the reference test grid."""
    from oracle import orc as O
    from test_oracle import GEOMETRY_SHAPES
    assert os.path.exists(BIN)
    text = "".join(f"{m} {k} {n}\n" for m, k, n in GEOMETRY_SHAPES)
    r = subprocess.run([BIN, "--geometry"], input=text, capture_output=True, text=True,
                       timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [tuple(int(x) for x in ln.split()) for ln in r.stdout.strip().splitlines()]
    assert [ln[:3] for ln in lines] == GEOMETRY_SHAPES
    for ln in lines:
        assert ln[3:] == O.alloc_geometry(*ln[:3])
        if O.ref_available():
            assert ln[3:] == O.ref_alloc_geometry(*ln[:3])
