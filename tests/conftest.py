import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gbs():
    """The HIP library, initialised on device 0.  Fails (does not skip) when
    the library or the GPU is missing: GPU tests must run the HIP path."""
    import gemmbitserial_amd as g
    g.init(0)
    return g


@pytest.fixture(scope="session")
def orc():
    from oracle import orc as o
    o.orc()
    return o


def mnist():
    d = np.load(os.path.join(ROOT, "tests", "golden", "mnist.npz"))
    n_out, n_in = (int(x) for x in d["shape"])
    bits = np.unpackbits(d["weights_bits"])[: n_out * n_in].reshape(n_out, n_in)
    w = bits.astype(np.int8) * 2 - 1
    return w, d["x"].copy(), d["golden"].copy()


def rand_values(rng, shape, nbits, issigned, dtype=np.int8):
    """Uniform values over the encoding's range (test/testhelpers.hpp:9-31):
    bipolar +-1 for 1-bit signed, [-2^(b-1), 2^(b-1)-1] signed,
    [0, 2^b - 1] unsigned."""
    if nbits == 1 and issigned:
        return np.where(rng.integers(0, 2, shape) == 1, 1, -1).astype(dtype)
    if issigned:
        return rng.integers(-(1 << (nbits - 1)), 1 << (nbits - 1), shape).astype(dtype)
    return rng.integers(0, 1 << nbits, shape).astype(dtype)
