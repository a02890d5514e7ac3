"""gbs_gemm_async_noalloc, the allocation-free GEMM entry point, through the
C ABI: its results equal gbs_gemm_async's (the existing entry point) and the
oracle's, bit for bit, on two calls in a row -- the second with smaller
operands, so stale scratch contents would show -- for the products whose
kernels use its scratch (bipolar x regular row sums on the popcount
engine), AUTO with and without value copies, and its size limit."""
import ctypes as C

import numpy as np
import pytest

from conftest import rand_values
from oracle import orc as O

pytestmark = pytest.mark.gpu

ENOTSUP, ERANGE = 5, 4


def _noalloc(gbs, l, r, d, algo, stream):
    return gbs.lib().gbs_gemm_async_noalloc(l.h, r.h, C.c_void_p(d.ptr), gbs.ALGOS[algo], stream)


def _pair(gbs, rng, M, K, N, lb, ls, rb, rs):
    a = rand_values(rng, (M, K), lb, ls)
    b = rand_values(rng, (N, K), rb, rs)
    want = O.gemm(O.import_regular(a, lb, ls), O.import_regular(b, rb, rs))
    return gbs.from_values(a, lb, ls), gbs.from_values(b, rb, rs), want


@pytest.mark.parametrize("bipolar_side", ["lhs", "rhs"])
def test_noalloc_popcount_row_sums_two_calls(gbs, orc, bipolar_side):
    """Bipolar x regular on the popcount engine (row sums in the scratch
    array), a large product then a smaller one, each against
    gbs_gemm_async and the oracle."""
    rng = np.random.default_rng(7 if bipolar_side == "lhs" else 8)
    stream = gbs.stream()
    ws = gbs.Workspace()
    for M, K, N in [(300, 777, 200), (90, 200, 40)]:
        if bipolar_side == "lhs":
            l, r, want = _pair(gbs, rng, M, K, N, 1, True, 3, True)
        else:
            l, r, want = _pair(gbs, rng, M, K, N, 3, False, 1, True)
        print(f"[noalloc] popcount {bipolar_side} bipolar {M}x{K}x{N}", flush=True)
        d_new, d_old = gbs.DeviceBuffer(4 * M * N), gbs.DeviceBuffer(4 * M * N)
        d_new.fill_zero()
        ws.reserve(l, r)
        assert _noalloc(gbs, l, r, d_new, "popcount", stream) == 0, gbs.lib().gbs_last_error()
        ws.gemm_async(l, r, d_old.ptr, "popcount", stream)
        gbs.sync(stream)
        got, old = d_new.to_host(np.int32, (N, M)), d_old.to_host(np.int32, (N, M))
        assert np.array_equal(got, old)
        assert np.array_equal(got, want)
        for x in (l, r, d_new, d_old):
            x.close()
    ws.close()


def test_noalloc_auto_without_and_with_copies(gbs, orc):
    """AUTO on fresh operands (no value copies: an engine that reads the
    planes) and after gbs_workspace_reserve made the copies (the matrix-core
    engine AUTO picks): both equal gbs_gemm_async's AUTO result; two
    products, the second smaller."""
    rng = np.random.default_rng(11)
    stream = gbs.stream()
    ws = gbs.Workspace()
    for M, K, N, lb, ls, rb, rs in [(512, 1500, 384, 2, True, 2, False),
                                    (130, 300, 70, 1, True, 2, False)]:
        l, r, want = _pair(gbs, rng, M, K, N, lb, ls, rb, rs)
        print(f"[noalloc] auto {M}x{K}x{N} {lb}{'s' if ls else 'u'} x {rb}{'s' if rs else 'u'}",
              flush=True)
        d_new, d_old = gbs.DeviceBuffer(4 * M * N), gbs.DeviceBuffer(4 * M * N)
        d_new.fill_zero()
        assert _noalloc(gbs, l, r, d_new, "auto", stream) == 0, gbs.lib().gbs_last_error()
        gbs.sync(stream)
        assert np.array_equal(d_new.to_host(np.int32, (N, M)), want)
        ws.reserve(l, r)
        d_new.fill_zero()
        assert _noalloc(gbs, l, r, d_new, "auto", stream) == 0, gbs.lib().gbs_last_error()
        ws.gemm_async(l, r, d_old.ptr, "auto", stream)
        gbs.sync(stream)
        got, old = d_new.to_host(np.int32, (N, M)), d_old.to_host(np.int32, (N, M))
        assert np.array_equal(got, old)
        assert np.array_equal(got, want)
        for x in (l, r, d_new, d_old):
            x.close()
    ws.close()


def test_noalloc_refuses_what_it_cannot_do(gbs, orc):
    """A matrix-core engine asked for explicitly without value copies, and
    a regular operand with more rows than the scratch holds, are refused
    with GBS_ERANGE before anything is launched."""
    lib = gbs.lib()
    rng = np.random.default_rng(3)
    stream = gbs.stream()
    l, r, _ = _pair(gbs, rng, 64, 128, 64, 2, False, 2, False)
    d = gbs.DeviceBuffer(4 * 64 * 64)
    assert _noalloc(gbs, l, r, d, "mfma", stream) == ERANGE
    assert b"copy" in lib.gbs_last_error()
    rows = 262145
    big = gbs.DeviceMatrix(2, rows, 64, False)
    bip = gbs.from_values(rand_values(rng, (16, 64), 1, True), 1, True)
    dbig = gbs.DeviceBuffer(4 * rows * 16)
    assert _noalloc(gbs, big, bip, dbig, "popcount", stream) == ERANGE
    assert b"scratch" in lib.gbs_last_error()
    gbs.sync(stream)
    for x in (l, r, d, big, bip, dbig):
        x.close()
