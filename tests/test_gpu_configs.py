"""Bit-exact parity of every engine at the benchmark configurations' full
sizes, on the benchmark's own seeded inputs: the plain GEMM (256 x 2048 x
256, 2u x 2s), the binary layers (4096 x 8192 x 512, bipolar and {0,1}) and
the conv-as-GEMM (256 x 2304 x 50176, 2s x 3u and 4s x 4u).  Each result is
compared with the oracle's gemm (oracle/gbs_oracle.c) on the oracle's own
packing of the same values, and with the reference compiled from its
sources (oracle/_ref) where that is built.  An engine that cannot hold a
product exactly must refuse it with GBS_ENOTSUP (and is then not AUTO's
choice).  The reference's own grid of shapes is test/test.cpp:110-155; the
dispatch being matched is arch-generic.hpp:331-345."""
import sys

import numpy as np
import pytest

from gemmbitserial_amd import workloads as W
from oracle import orc as O

pytestmark = pytest.mark.gpu

ENGINES = ["auto", "popcount", "mfma", "mfma_fp4", "mfma_fp6", "mfma_bits", "mfma_planes"]
CONFIGS = {"plain_gemm": (0, W.plain_gemm), "binary_layers": (1, W.binary_layers),
           "conv_gemm": (2, W.conv_gemm)}


def _cases():
    out = []
    for name, (cid, make) in CONFIGS.items():
        n_runs = 2 if name != "plain_gemm" else 1
        for i in range(n_runs):
            out.append((name, i))
    return out


_cache = {}


def _run_inputs(name, i):
    """(encoding, lhs, rhs, lb, rb, ls, rs) of run i of a configuration, on
    np.random.default_rng(SEED0 + config number) exactly as bench.py draws
    them, plus the oracle's result (computed once)."""
    key = (name, i)
    if key not in _cache:
        cid, make = CONFIGS[name]
        runs = make(np.random.default_rng(W.SEED0 + cid))
        for j, run in enumerate(runs):
            enc, a, b, lb, rb, ls, rs = run
            _cache[(name, j)] = (run, None)
    run, want = _cache[key]
    if want is None:
        enc, a, b, lb, rb, ls, rs = run
        want = O.gemm(O.import_regular(a, lb, ls), O.import_regular(b, rb, rs))
        _cache[key] = (run, want)
    return run, want


@pytest.mark.parametrize("name,i", _cases())
def test_config_full_size_every_engine(gbs, orc, name, i):
    (enc, a, b, lb, rb, ls, rs), want = _run_inputs(name, i)
    M, K = a.shape
    N = b.shape[0]
    print(f"[configs] {name} {enc}: {M}x{K}x{N}, oracle done", flush=True)
    l = gbs.from_values(a, lb, ls)
    r = gbs.from_values(b, rb, rs)
    # the device planes are the oracle's planes, padding included
    assert np.array_equal(l.planes(), O.import_regular(a, lb, ls, nrows_a=l.nrows_a,
                                                       ncols_a=l.ncols_a).planes)
    auto = gbs.select(l, r)
    exact = []
    for eng in ENGINES:
        print(f"[configs] {name} {enc}: engine {eng}", flush=True)
        try:
            got = gbs.gemm(l, r, eng)
        except gbs.GbsError as e:
            assert e.code == 5, str(e)          # GBS_ENOTSUP: cannot hold it exactly
            assert eng not in ("auto", auto, "popcount")
            continue
        assert got.shape == (N, M)
        assert np.array_equal(got, want), f"{name} {enc}: engine {eng} differs from the oracle"
        exact.append(eng)
    assert auto in exact and "popcount" in exact
    if O.ref_available():
        ref = O.ref_gemm(a, b, lb, rb, ls, rs)
        assert np.array_equal(ref, want), "the compiled reference differs from the oracle"
    l.close()
    r.close()
    sys.stdout.flush()


def _bits_case(gbs, M, K, N, bip_l, bip_r, seed):
    rng = np.random.default_rng(seed)
    a = (np.where(rng.random((M, K)) < 0.5, 1, -1) if bip_l
         else (rng.random((M, K)) < 0.3)).astype(np.int8)
    b = (np.where(rng.random((N, K)) < 0.5, 1, -1) if bip_r
         else (rng.random((N, K)) < 0.3)).astype(np.int8)
    l = gbs.from_values(a, 1, bip_l)
    r = gbs.from_values(b, 1, bip_r)
    want = O.gemm(O.import_regular(a, 1, bip_l), O.import_regular(b, 1, bip_r))
    got = gbs.gemm(l, r, "mfma_bits")
    l.close()
    r.close()
    return got, want


@pytest.mark.parametrize("M,K,N", [(256, 2500, 128), (512, 3072, 64), (128, 2049, 256),
                                   (256, 1000, 128), (64, 2048, 512), (1024, 2560, 512)])
@pytest.mark.parametrize("bip", [(True, True), (False, False), (True, False), (False, True)])
def test_bits_engine_xcd_order_and_three_super_steps(gbs, orc, M, K, N, bip):
    """1-bit matrix-core engine on shapes whose 64 x 64 tile count is a
    multiple of 8 (the XCD-ordered tile mapping) and depths of three
    super-steps (2049..3072) as well as the ring kernel's (<= 2048)."""
    tiles = ((M + 63) // 64) * ((N + 63) // 64)
    assert tiles % 8 == 0
    print(f"[configs] bits {M}x{K}x{N} bipolar {bip}", flush=True)
    got, want = _bits_case(gbs, M, K, N, bip[0], bip[1], M + K + N)
    assert np.array_equal(got, want)
