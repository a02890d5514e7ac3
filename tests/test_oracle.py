"""Pins the CPU oracle (oracle/gbs_oracle.c) before it is trusted:
against the reference compiled from its own sources (oracle/_ref), against
independent naive products, against the MNIST known-answer vector the
reference's tests hold, and against closed-form identities."""
import itertools
import os

import numpy as np
import pytest

from conftest import ROOT, mnist, rand_values
from oracle import orc as O

needs_ref = pytest.mark.skipif(not O.ref_available(),
                               reason="oracle/_ref not built (make -C oracle ref)")

DIMS = [3, 5, 7, 16, 17, 18, 30, 31, 32, 100, 177]  # test.cpp:112


# ---------------------------------------------------------------- geometry
@needs_ref
@pytest.mark.parametrize("M,K,N", [(256, 2048, 256), (4096, 8192, 512), (256, 2304, 50176),
                                   (256, 784, 1), (3, 6, 9), (17, 34, 1), (1000, 4096, 1),
                                   (96, 363, 3025), (4096, 9216, 1), (100, 200, 300)])
def test_alloc_geometry_matches_reference(orc, M, K, N):
    assert O.alloc_geometry(M, K, N) == O.ref_alloc_geometry(M, K, N)


def test_alloc_geometry_closed_form(orc):
    # N == 1: register blocking (1,4,1): depth aligned to 256, all rows in one block
    lb, rb, lra, lca, rra, rca = O.alloc_geometry(256, 784, 1)
    assert (lca, rca, lra, rra) == (1024, 1024, 256, 1)
    # rows always a multiple of the block, depth of 64
    for M, K, N in [(4096, 8192, 512), (37, 100, 59)]:
        lb, rb, lra, lca, rra, rca = O.alloc_geometry(M, K, N)
        assert lra % lb == 0 and rra % rb == 0 and lca % 64 == 0 and lra >= M


# ------------------------------------------------------------------ import
@needs_ref
@pytest.mark.parametrize("nbits,issigned", [(1, True), (1, False), (2, True), (2, False),
                                            (3, True), (4, False), (7, True), (8, False)])
@pytest.mark.parametrize("col_major", [False, True])
def test_import_words_match_reference(orc, nbits, issigned, col_major):
    rng = np.random.default_rng(100 + nbits)
    for (r, c) in [(4, 4), (17, 77), (33, 130), (5, 1023)]:
        # out-of-range values exercise the low-bit truncation; zeros in
        # bipolar become -1
        vals = rng.integers(-128, 128, (r, c)).astype(np.int8)
        src = vals.T.copy() if col_major else vals
        ref = O.ref_import_regular(src, nbits, issigned, col_major=col_major)
        mine = O.import_regular(src, nbits, issigned, col_major=col_major)
        assert np.array_equal(ref.planes, mine.planes)


@needs_ref
@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16, np.int32,
                                   np.uint32, np.int64, np.uint64])
@pytest.mark.parametrize("nbits,issigned", [(1, True), (3, True), (3, False)])
def test_import_dtypes_match_reference(orc, dtype, nbits, issigned):
    rng = np.random.default_rng(7)
    info = np.iinfo(dtype)
    lo, hi = max(info.min, -1000), min(info.max, 1000)
    vals = rng.integers(lo, hi, (9, 70)).astype(dtype)
    assert np.array_equal(O.ref_import_regular(vals, nbits, issigned).planes,
                          O.import_regular(vals, nbits, issigned).planes)


@needs_ref
@pytest.mark.parametrize("nbits,issigned", [(1, True), (2, False), (3, True)])
def test_import_raw_uint8_overload_matches_reference(orc, nbits, issigned):
    vals = np.random.default_rng(3).integers(0, 256, (11, 90)).astype(np.uint8)
    ref = O.ref_import_regular(vals, nbits, issigned, raw_u8=True)
    assert np.array_equal(ref.planes, O.import_regular(vals, nbits, issigned, raw_u8=True).planes)
    if nbits == 1 and issigned:
        # raw overload takes bit 0, the template takes "> 0": they differ
        assert not np.array_equal(ref.planes, O.import_regular(vals, 1, True).planes)


def test_import_padding_is_zero(orc):
    vals = np.full((5, 70), -1, np.int8)
    p = O.import_regular(vals, 2, True, nrows_a=8, ncols_a=192)
    planes = p.planes.reshape(2, 8, 3)
    assert (planes[:, 5:, :] == 0).all() and (planes[:, :5, 2] == 0).all()
    assert (planes[:, :5, 1] == (1 << 6) - 1).all() and (planes[:, :5, 0] == 2**64 - 1).all()


@pytest.mark.parametrize("nbits,issigned", [(1, True), (1, False), (2, True), (3, False),
                                            (4, True), (7, True), (8, False)])
def test_export_round_trip(orc, nbits, issigned):
    rng = np.random.default_rng(nbits)
    for d in [4, 17, 77, 100]:
        vals = rand_values(rng, (d, d), nbits, issigned, np.int16)
        if nbits == 8 and not issigned:
            vals = vals.astype(np.uint8)
        p = O.import_regular(vals, nbits, issigned)
        assert np.array_equal(O.export_regular(p), vals.astype(np.int32))


@needs_ref
def test_export_and_sum_rows_match_reference(orc):
    rng = np.random.default_rng(11)
    for nbits, issigned in itertools.product([1, 2, 3, 4], [True, False]):
        for d in [4, 17, 77, 1023]:
            vals = rng.integers(-100, 100, (d, d)).astype(np.int8)
            p = O.import_regular(vals, nbits, issigned)
            assert np.array_equal(O.export_regular(p), O.ref_export_regular(p))
            assert np.array_equal(O.sum_rows(p), O.ref_sum_rows(p))


def test_sum_rows_closed_form(orc):
    rng = np.random.default_rng(5)
    for nbits, issigned in itertools.product([1, 2, 3, 4], [True, False]):
        vals = rand_values(rng, (31, 1023), nbits, issigned)
        p = O.import_regular(vals, nbits, issigned)
        assert np.array_equal(O.sum_rows(p), vals.astype(np.int64).sum(1))


def threshold_inputs(rng, dtype, nrows, ncols, nthres):
    """Values and ascending per-row thresholds of a dtype, spread over the
    type's range so that comparisons differ from int32-widened ones (uint32
    and uint64 values above the signed maximum, negative and fractional
    floats)."""
    dt = np.dtype(dtype)
    if dt.kind == "f":
        vals = rng.uniform(-3.0, 300.0, (nrows, ncols)).astype(dt)
        vals[rng.random((nrows, ncols)) < 0.1] = 0.5
        thr = np.sort(rng.uniform(-3.0, 300.0, (nthres, nrows)), axis=0).astype(dt)
        return vals, thr
    info = np.iinfo(dt)
    lo, hi = int(info.min), int(info.max)
    if dt.kind == "u" and dt.itemsize >= 4:
        lo = hi // 2 - 4096        # straddle 2^31 (2^63): int32 widening flips these
        hi = hi // 2 + 4096
    vals = rng.integers(lo, hi, (nrows, ncols), dtype=dt, endpoint=True)
    thr = np.sort(rng.integers(lo, hi, (nthres, nrows), dtype=dt, endpoint=True), axis=0)
    if nthres:
        vals[:, 0] = thr[0]          # equal to a threshold: "<=" takes it
    return vals, thr


QUANT_DTYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64,
                np.uint64, np.float32, np.float64]


@pytest.mark.parametrize("dtype", QUANT_DTYPES, ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("nthres", [0, 1, 3, 5])
def test_import_and_quantize_matches_reference(orc, dtype, nthres):
    """Word for word against the compiled reference's importRegularAndQuantize<T>
    (gemmbitserial.hpp:281-308), row- and column-major, ragged shapes."""
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(nthres * 31 + np.dtype(dtype).itemsize)
    for nrows, ncols in [(6, 90), (17, 129), (1, 64), (33, 3)]:
        vals, thr = threshold_inputs(rng, dtype, nrows, ncols, nthres)
        for col_major in (False, True):
            src = np.ascontiguousarray(vals.T) if col_major else vals
            for nbits in (1, 3, 8):
                got = O.import_and_quantize(src, thr, nbits, col_major=col_major)
                want = O.ref_import_and_quantize(src, thr, nbits, col_major=col_major)
                assert (got.nrows_a, got.ncols_a) == (want.nrows_a, want.ncols_a)
                assert np.array_equal(got.planes, want.planes), (nrows, ncols, col_major, nbits)


@pytest.mark.parametrize("dtype", QUANT_DTYPES, ids=lambda d: np.dtype(d).name)
def test_import_and_quantize_closed_form(orc, dtype):
    """With ascending thresholds the code is the number of thresholds below
    the element (searchsorted, left side); with none it is the element
    itself modulo 256 (truncated toward zero for floats)."""
    rng = np.random.default_rng(np.dtype(dtype).itemsize * 7)
    for nthres in (0, 1, 3, 7):
        vals, thr = threshold_inputs(rng, dtype, 9, 150, nthres)
        p = O.import_and_quantize(vals, thr, 8)
        got = O.export_regular(p)
        if nthres:
            want = np.array([[np.searchsorted(thr[:, r], vals[r, c], side="left")
                              for c in range(vals.shape[1])] for r in range(vals.shape[0])])
        elif np.dtype(dtype).kind == "f":
            want = np.trunc(vals.astype(np.float64)).astype(np.int64) % 256
        else:
            want = vals.astype(object) % 256
        assert np.array_equal(got, np.asarray(want, dtype=np.int64))


def test_import_and_quantize_catches_int32_widening(orc):
    """A planted mistake -- comparing uint32 values as int32 -- changes the
    planes, so the reference comparison above would catch it."""
    vals = np.array([[0x80000005, 7, 0xfffffff0]], np.uint32)
    thr = np.array([[0x80000000], [0x90000000]], np.uint32)
    good = O.export_regular(O.import_and_quantize(vals, thr, 2))
    widened = O.export_regular(O.import_and_quantize(vals.view(np.int32), thr.view(np.int32), 2))
    assert good.tolist() == [[1, 0, 2]]
    assert not np.array_equal(good, widened)


# ------------------------------------------------------------ geometry probe
GEOMETRY_SHAPES = [(256, 2048, 256), (4096, 8192, 512), (256, 2304, 50176), (256, 784, 1),
                   (64, 256, 1), (4096, 8192, 256), (1089, 1808, 64), (3886, 2374, 1),
                   (3, 6, 9), (5, 10, 15), (17, 34, 51), (177, 354, 531), (256, 512, 768),
                   (100, 200, 1), (2, 3, 4), (1, 1, 1), (9216, 4096, 1), (96, 363, 3025)]


def test_header_geometry_matches_reference(orc):
    """The drop-in header's allocGEMMContext geometry (lhsBlock, rhsBlock,
    nrows_a, ncols_a of both sides), computed by tests/cpp/geometry_probe.cpp
    from include/gemmbitserial.hpp, equals the compiled reference's."""
    import subprocess
    probe = os.path.join(ROOT, "gemmbitserial_amd", "bin", "geometry_probe")
    assert os.path.exists(probe), "build the C++ programs first (python -m gemmbitserial_amd.build)"
    text = "".join(f"{m} {k} {n}\n" for m, k, n in GEOMETRY_SHAPES)
    r = subprocess.run([probe], input=text, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = [tuple(int(x) for x in ln.split()) for ln in r.stdout.strip().splitlines()]
    assert len(lines) == len(GEOMETRY_SHAPES)
    for (m, k, n), got in zip(GEOMETRY_SHAPES, lines):
        assert got[:3] == (m, k, n)
        assert got[3:] == O.alloc_geometry(m, k, n), (m, k, n)
        if O.ref_available():
            assert got[3:] == O.ref_alloc_geometry(m, k, n), (m, k, n)


# -------------------------------------------------------------------- gemm
def _check_gemm(lhs_vals, rhs_vals, lb, rb, ls, rs, use_ref):
    lp = O.import_regular(lhs_vals, lb, ls)
    rp = O.import_regular(rhs_vals, rb, rs)
    got = O.gemm(lp, rp)
    lv = O.values_of(lhs_vals, lb, ls)
    rv = O.values_of(rhs_vals, rb, rs)
    assert np.array_equal(got, O.wrap_i32(O.naive_gemm_i64(lv, rv)))
    assert np.array_equal(got, O.wrap_i32(O.blas_gemm_exact(lv, rv)))
    if use_ref:
        assert np.array_equal(got, O.ref_gemm(lhs_vals, rhs_vals, lb, rb, ls, rs))


@pytest.mark.parametrize("lb,rb", list(itertools.product([1, 2, 3, 4], repeat=2)))
@pytest.mark.parametrize("ls,rs", list(itertools.product([False, True], repeat=2)))
def test_gemm_all_encodings(orc, lb, rb, ls, rs):
    rng = np.random.default_rng(lb * 100 + rb * 10 + 2 * ls + rs)
    for (M, K, N) in [(3, 6, 9), (17, 34, 1), (31, 62, 93), (18, 200, 5)]:
        lhs = rand_values(rng, (M, K), lb, ls)
        rhs = rand_values(rng, (N, K), rb, rs)
        _check_gemm(lhs, rhs, lb, rb, ls, rs, O.ref_available())


def test_gemm_reference_test_grids(orc):
    # test_matrix_matrix (test.cpp:110-155): unsigned lb = rb in {2,3,4}, M=d, K=2d, N in {1, 3d}
    rng = np.random.default_rng(1)
    for b, d, mv in itertools.product([2, 3, 4], DIMS, [0, 1]):
        lhs = rand_values(rng, (d, 2 * d), b, False)
        rhs = rand_values(rng, (1 if mv else 3 * d, 2 * d), b, False)
        _check_gemm(lhs, rhs, b, b, False, False, O.ref_available() and d < 100)


def test_gemm_bipolar_identity(orc):
    rng = np.random.default_rng(9)
    a = rand_values(rng, (40, 300), 1, True)
    b = rand_values(rng, (25, 300), 1, True)
    pa, pb = O.import_regular(a, 1, True), O.import_regular(b, 1, True)
    # a.b = K - 2 popcount(a ^ b) computed on the packed words directly
    wa = pa.planes.reshape(40, -1)
    wb = pb.planes.reshape(25, -1)
    pc = np.array([[sum(bin(int(x) ^ int(y)).count("1") for x, y in zip(wa[m], wb[n]))
                    for m in range(40)] for n in range(25)])
    assert np.array_equal(O.gemm(pa, pb), 300 - 2 * pc)


def test_gemm_wraps_like_int32(orc):
    # 8-bit x 8-bit unsigned, K large enough to exceed 2^31
    K = 40000
    lhs = np.full((2, K), 255, np.uint8)
    rhs = np.full((3, K), 255, np.uint8)
    lp, rp = O.import_regular(lhs, 8, False), O.import_regular(rhs, 8, False)
    assert np.array_equal(O.gemm(lp, rp), O.wrap_i32(np.full((3, 2), 255 * 255 * K)))


def test_mnist_known_answer(orc):
    w, x, golden = mnist()
    lp = O.import_regular(w, 1, True)
    rp = O.import_regular(x.reshape(1, -1), 2, False)
    assert np.array_equal(O.gemm(lp, rp).reshape(-1), golden)
    if O.ref_available():
        assert np.array_equal(O.ref_gemm(w, x.reshape(1, -1), 1, 2, True, False).reshape(-1),
                              golden)


def test_planted_mistake_is_caught(orc):
    """A dropped MSB sign in the oracle's input would fail the naive check."""
    rng = np.random.default_rng(4)
    lhs = rand_values(rng, (9, 70), 3, True)
    rhs = rand_values(rng, (7, 70), 2, True)
    lp = O.import_regular(lhs, 3, True)
    lp_wrong = O.Planes(lp.planes, 3, False, lp.nrows, lp.ncols, lp.nrows_a, lp.ncols_a)
    rp = O.import_regular(rhs, 2, True)
    want = O.wrap_i32(O.naive_gemm_i64(lhs.astype(np.int32), rhs.astype(np.int32)))
    assert np.array_equal(O.gemm(lp, rp), want)
    assert not np.array_equal(O.gemm(lp_wrong, rp), want)
