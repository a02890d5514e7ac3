"""GPU parity: every HIP kernel against the oracle (and, where built, the
reference compiled from source) on the same seeded inputs.  Bit-exact."""
import itertools
import sys

import numpy as np
import pytest

from conftest import mnist, rand_values
from oracle import orc as O


def align(x, a):
    return (x + a - 1) // a * a

pytestmark = pytest.mark.gpu
DIMS = [3, 5, 7, 16, 17, 18, 30, 31, 32, 100, 177, 256]  # test.cpp:112


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------------ import
@pytest.mark.parametrize("nbits,issigned", [(1, True), (1, False), (2, True), (2, False),
                                            (3, True), (4, False), (7, True), (8, False)])
@pytest.mark.parametrize("col_major", [False, True])
def test_pack_words_match_oracle(gbs, orc, nbits, issigned, col_major):
    rng = np.random.default_rng(200 + nbits)
    for (r, c, ra, ca) in [(4, 4, 4, 64), (17, 77, 18, 128), (33, 130, 40, 192),
                           (5, 1023, 5, 1024), (64, 4096, 70, 4096), (1, 3, 2, 256)]:
        say("pack", nbits, issigned, col_major, r, c)
        vals = rng.integers(-128, 128, (r, c)).astype(np.int8)
        src = vals.T.copy() if col_major else vals
        want = O.import_regular(src, nbits, issigned, nrows_a=ra, ncols_a=ca,
                                col_major=col_major)
        m = gbs.DeviceMatrix(nbits, r, c, issigned, ra, ca)
        got = m.import_regular(src, col_major=col_major, want_planes=True)
        assert np.array_equal(got, want.planes)
        assert np.array_equal(m.planes(), want.planes)


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16, np.int32,
                                   np.uint32, np.int64, np.uint64])
def test_pack_dtypes(gbs, orc, dtype):
    rng = np.random.default_rng(8)
    info = np.iinfo(dtype)
    vals = rng.integers(max(info.min, -1000), min(info.max, 1000), (9, 70)).astype(dtype)
    for nbits, issigned in [(1, True), (3, True), (3, False), (8, True)]:
        say("dtype", dtype, nbits, issigned)
        m = gbs.DeviceMatrix(nbits, 9, 70, issigned)
        got = m.import_regular(vals, want_planes=True)
        assert np.array_equal(got, O.import_regular(vals, nbits, issigned).planes)


@pytest.mark.parametrize("dtype", [np.int8, np.uint8])
def test_pack_byte_rows_vector_path(gbs, orc, dtype):
    """Rows of one-byte elements that start 16-byte aligned and hold a
    multiple of 16 bytes take the vector path (16 codes per thread, bits
    gathered by u8 dot products): every encoding, the raw overload, zeros,
    extremes and a ragged last quarter in a second matrix."""
    rng = np.random.default_rng(77)
    info = np.iinfo(dtype)
    for (r, c) in [(33, 256), (7, 1024 + 16)]:
        vals = rng.integers(info.min, info.max + 1, (r, c)).astype(dtype)
        vals[0, :16] = 0
        vals[1, :16] = info.max
        vals[2, :16] = info.min
        for nbits, issigned in [(1, True), (1, False), (2, True), (3, False), (5, True), (8, False)]:
            say("byte rows", dtype, r, c, nbits, issigned)
            m = gbs.DeviceMatrix(nbits, r, c, issigned)
            got = m.import_regular(vals, want_planes=True)
            assert np.array_equal(got, O.import_regular(vals, nbits, issigned).planes)
            if dtype == np.uint8:
                got = m.import_regular(vals, raw_u8=True, want_planes=True)
                assert np.array_equal(got, O.import_regular(vals, nbits, issigned,
                                                            raw_u8=True).planes)


def test_pack_raw_uint8(gbs, orc):
    vals = np.random.default_rng(3).integers(0, 256, (11, 90)).astype(np.uint8)
    for nbits, issigned in [(1, True), (2, False), (3, True)]:
        m = gbs.DeviceMatrix(nbits, 11, 90, issigned)
        got = m.import_regular(vals, raw_u8=True, want_planes=True)
        assert np.array_equal(got, O.import_regular(vals, nbits, issigned, raw_u8=True).planes)


def test_pack_overwrites_previous_content(gbs, orc):
    m = gbs.DeviceMatrix(2, 7, 100, True, 8, 128)
    m.import_regular(np.full((7, 100), -1, np.int8))
    vals = np.random.default_rng(1).integers(-2, 2, (7, 100)).astype(np.int8)
    got = m.import_regular(vals, want_planes=True)
    assert np.array_equal(got, O.import_regular(vals, 2, True, nrows_a=8, ncols_a=128).planes)


QUANT_DTYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64,
                np.uint64, np.float32, np.float64]


@pytest.mark.parametrize("dtype", QUANT_DTYPES, ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("nthres", [0, 1, 3])
def test_import_quantize(gbs, orc, dtype, nthres):
    """importRegularAndQuantize<T> on the GPU equals the oracle word for
    word (the oracle equals the compiled reference, tests/test_oracle.py):
    comparisons in T (uint32 / uint64 values straddling 2^31 / 2^63,
    floats), no thresholds (raw low bits), row- and column-major, ragged
    shapes, padding rows."""
    from test_oracle import threshold_inputs
    rng = np.random.default_rng(nthres * 13 + np.dtype(dtype).itemsize)
    for nrows, ncols in [(6, 90), (37, 200), (1, 64), (70, 3)]:
        vals, thr = threshold_inputs(rng, dtype, nrows, ncols, nthres)
        for col_major in (False, True):
            src = np.ascontiguousarray(vals.T) if col_major else vals
            for nbits in (2, 8):
                want = O.import_and_quantize(src, thr, nbits, nrows_a=nrows + 3,
                                             ncols_a=align(ncols, 128), col_major=col_major)
                m = gbs.DeviceMatrix(nbits, nrows, ncols, False, nrows + 3, align(ncols, 128))
                got = m.import_quantize(src, thr, col_major=col_major, want_planes=True)
                m.close()
                assert np.array_equal(got, want.planes), (nrows, ncols, col_major, nbits)


@pytest.mark.parametrize("nbits,issigned", list(itertools.product([1, 2, 3, 4, 8], [True, False])))
def test_export_and_sum_rows(gbs, orc, nbits, issigned):
    rng = np.random.default_rng(nbits)
    for d in [4, 17, 77, 1023]:
        vals = rng.integers(-100, 100, (d, d)).astype(np.int8)
        m = gbs.from_values(vals, nbits, issigned)
        want = O.import_regular(vals, nbits, issigned)
        assert np.array_equal(m.export_regular(), O.export_regular(want))
        assert np.array_equal(m.sum_rows(), O.sum_rows(want))


# --------------------------------------------------------------------- gemm
def run_gemm(gbs, lhs, rhs, lb, rb, ls, rs, algo="popcount", geom=None):
    M, K = lhs.shape
    N = rhs.shape[0]
    if geom is None:
        geom = O.alloc_geometry(M, K, N)
    l = gbs.DeviceMatrix(lb, M, K, ls, geom[2], geom[3])
    r = gbs.DeviceMatrix(rb, N, K, rs, geom[4], geom[5])
    l.import_regular(lhs)
    r.import_regular(rhs)
    return gbs.gemm(l, r, algo)


def expected(lhs, rhs, lb, rb, ls, rs):
    lp = O.import_regular(lhs, lb, ls)
    rp = O.import_regular(rhs, rb, rs)
    return O.gemm(lp, rp)


@pytest.mark.parametrize("lb,rb", list(itertools.product([1, 2, 3, 4], repeat=2)))
@pytest.mark.parametrize("ls,rs", list(itertools.product([False, True], repeat=2)))
def test_gemm_all_encodings(gbs, orc, lb, rb, ls, rs):
    rng = np.random.default_rng(lb * 100 + rb * 10 + 2 * ls + rs)
    for (M, K, N) in [(3, 6, 9), (17, 34, 1), (31, 62, 93), (18, 200, 5), (2, 70, 3),
                      (130, 700, 67), (64, 1030, 200)]:
        say("gemm", lb, rb, ls, rs, M, K, N)
        lhs = rand_values(rng, (M, K), lb, ls)
        rhs = rand_values(rng, (N, K), rb, rs)
        got = run_gemm(gbs, lhs, rhs, lb, rb, ls, rs)
        assert np.array_equal(got, expected(lhs, rhs, lb, rb, ls, rs))


def test_gemm_reference_grids(gbs, orc):
    """test_matrix_matrix, test_bipolar_times_regular, test_bipolar_times_bipolar
    (test.cpp:110-277) at the reference's dimension list."""
    rng = np.random.default_rng(1)
    for b, d, mv in itertools.product([2, 3, 4], DIMS, [0, 1]):
        lhs = rand_values(rng, (d, 2 * d), b, False)
        rhs = rand_values(rng, (1 if mv else 3 * d, 2 * d), b, False)
        assert np.array_equal(run_gemm(gbs, lhs, rhs, b, b, False, False),
                              O.wrap_i32(O.naive_gemm_i64(lhs, rhs)))
    for sw, rbits, d, sg in itertools.product([0, 1], [2, 3, 4], DIMS, [True, False]):
        bip = rand_values(rng, (d, d), 1, True)
        reg = rand_values(rng, (d, d), rbits, sg)
        if sw:
            got = run_gemm(gbs, reg, bip, rbits, 1, sg, True)
            want = O.naive_gemm_i64(reg, bip)
        else:
            got = run_gemm(gbs, bip, reg, 1, rbits, True, sg)
            want = O.naive_gemm_i64(bip, reg)
        assert np.array_equal(got, O.wrap_i32(want))
    for d, mv in itertools.product(DIMS, [0, 1]):
        lhs = rand_values(rng, (d, d), 1, True)
        rhs = rand_values(rng, (1 if mv else d, d), 1, True)
        assert np.array_equal(run_gemm(gbs, lhs, rhs, 1, 1, True, True),
                              O.wrap_i32(O.naive_gemm_i64(lhs, rhs)))


@pytest.mark.parametrize("lb,rb", [(5, 3), (8, 8), (6, 1), (1, 7)])
def test_gemm_wide_bitwidths(gbs, orc, lb, rb):
    rng = np.random.default_rng(lb + rb)
    for ls, rs in itertools.product([False, True], repeat=2):
        lhs = rand_values(rng, (37, 300), lb, ls, np.int16)
        rhs = rand_values(rng, (45, 300), rb, rs, np.int16)
        got = run_gemm(gbs, lhs, rhs, lb, rb, ls, rs)
        assert np.array_equal(got, expected(lhs, rhs, lb, rb, ls, rs))


def test_gemm_wraps_like_int32(gbs, orc):
    K = 40000
    lhs = np.full((2, K), 255, np.uint8)
    rhs = np.full((9, K), 255, np.uint8)
    got = run_gemm(gbs, lhs, rhs, 8, 8, False, False)
    assert np.array_equal(got, O.wrap_i32(np.full((9, 2), 255 * 255 * K)))


def test_gemm_empty_and_degenerate(gbs, orc):
    for (M, K, N) in [(5, 0, 3), (0, 10, 3), (4, 10, 0), (1, 1, 1)]:
        lhs = np.ones((M, K), np.int8)
        rhs = np.ones((N, K), np.int8)
        l = gbs.DeviceMatrix(2, M, K, False)
        r = gbs.DeviceMatrix(2, N, K, False)
        if M * K:
            l.import_regular(lhs)
        if N * K:
            r.import_regular(rhs)
        got = gbs.gemm(l, r, "popcount")
        assert got.shape == (N, M) and np.array_equal(got, np.full((N, M), K, np.int32))


def test_gemm_depth_mismatch_is_an_error(gbs):
    l = gbs.DeviceMatrix(2, 4, 10, False)
    r = gbs.DeviceMatrix(2, 4, 11, False)
    with pytest.raises(gbs.GbsError, match="depths must match"):
        gbs.gemm(l, r)


def test_mnist_known_answer(gbs, orc):
    w, x, golden = mnist()
    geom = O.alloc_geometry(256, 784, 1)
    got = run_gemm(gbs, w, x.reshape(1, -1), 1, 2, True, False, geom=geom)
    assert np.array_equal(got.reshape(-1), golden)


def test_split_k_path(gbs, orc):
    # few tiles, deep K: exercises the K-split atomic path
    rng = np.random.default_rng(12)
    for lb, rb, ls, rs in [(1, 1, False, False), (1, 1, True, True), (2, 3, True, False),
                           (1, 2, True, False)]:
        lhs = rand_values(rng, (100, 20000), lb, ls)
        rhs = rand_values(rng, (40, 20000), rb, rs)
        got = run_gemm(gbs, lhs, rhs, lb, rb, ls, rs)
        lv, rv = O.values_of(lhs, lb, ls), O.values_of(rhs, rb, rs)
        assert np.array_equal(got, O.wrap_i32(O.blas_gemm_exact(lv, rv)))


ENCODINGS = [(1, True), (1, False), (2, True), (2, False), (3, True), (3, False),
             (4, True), (4, False)]


@pytest.mark.parametrize("lenc", ENCODINGS)
def test_skinny_engine_every_encoding(gbs, orc, lenc):
    """Products of 1..8 rhs rows go to the skinny kernel (planes straight
    from HBM, rhs pieces in LDS, plane pairs folded as they are counted,
    bipolar sides through a ones plane): every encoding pair, depths
    around word, 128-bit and 8192-column piece boundaries, one to a few
    thousand rows."""
    lb, ls = lenc
    rng = np.random.default_rng(900 + 2 * lb + ls)
    shapes = [(1, 1, 1), (7, 63, 2), (100, 65, 3), (1000, 200, 8), (513, 8192, 5),
              (300, 8193, 4), (64, 20000, 7), (2049, 1000, 1), (33, 128, 6)]
    for (rb, rs) in ENCODINGS:
        for (M, K, N) in shapes:
            say("skinny", lb, ls, rb, rs, M, K, N)
            lhs = rand_values(rng, (M, K), lb, ls)
            rhs = rand_values(rng, (N, K), rb, rs)
            got = run_gemm(gbs, lhs, rhs, lb, rb, ls, rs)
            want = O.wrap_i32(O.blas_gemm_exact(O.values_of(lhs, lb, ls),
                                                O.values_of(rhs, rb, rs)))
            assert np.array_equal(got, want)


def test_skinny_engine_fill_and_fallback(gbs, orc):
    """Many rows (the plan widens the lanes per row to fill the chip) and
    rows of an odd number of words (the last 128-bit slice of a row holds
    one real word; the word after it belongs to the next row and must not
    be counted)."""
    rng = np.random.default_rng(951)
    for (lb, ls, rb, rs) in [(1, True, 1, True), (1, False, 1, False), (2, True, 4, False),
                             (1, True, 3, True), (4, True, 1, True)]:
        for N in (1, 2, 8):
            say("skinny fill", lb, ls, rb, rs, N)
            lhs = rand_values(rng, (12000, 2048), lb, ls)
            rhs = rand_values(rng, (N, 2048), rb, rs)
            got = run_gemm(gbs, lhs, rhs, lb, rb, ls, rs)
            assert np.array_equal(got, O.wrap_i32(O.blas_gemm_exact(
                O.values_of(lhs, lb, ls), O.values_of(rhs, rb, rs))))
    for (lb, ls, rb, rs) in [(2, True, 2, False), (1, True, 2, False), (1, True, 1, True),
                             (3, False, 1, True), (4, True, 4, True)]:
        for (M, K, N, ka) in [(50, 150, 3, 192), (77, 8190, 8, 8256), (40, 9000, 2, 9024)]:
            say("skinny odd pitch", lb, ls, rb, rs, M, K, N, ka)
            lhs = rand_values(rng, (M, K), lb, ls)
            rhs = rand_values(rng, (N, K), rb, rs)
            got = run_gemm(gbs, lhs, rhs, lb, rb, ls, rs, geom=(M, N, M, ka, N, ka))
            assert np.array_equal(got, O.wrap_i32(O.naive_gemm_i64(lhs, rhs)))


# ------------------------------------------- 1-bit matrix-core engine on planes
@pytest.mark.parametrize("ls,rs", list(itertools.product([False, True], repeat=2)))
def test_bits_engine_every_1bit_encoding(gbs, orc, ls, rs):
    """The matrix-core engine that reads 1-bit operands straight from their
    planes (bits turned into FP4 codes in registers): bipolar and {0,1} on
    either side; ragged tiles of both operands; depths around word,
    1024-bit step and four-step boundaries (bipolar x bipolar padding is
    subtracted); rows of an odd number of words; one result per engine."""
    rng = np.random.default_rng(1300 + 2 * ls + rs)
    shapes = [(1, 1, 1), (3, 6, 9), (17, 63, 1), (64, 64, 64), (65, 65, 63),
              (130, 1000, 67), (64, 1024, 200), (200, 1025, 64), (129, 4095, 130),
              (257, 4097, 70), (100, 8192, 33), (70, 20000, 150),
              # tile counts divisible by 8 (the XCD-grouped tile order) and
              # three super-steps (the register form's first depths)
              (256, 2500, 128), (512, 3072, 64), (130, 2049, 520), (1024, 2112, 512)]
    for (M, K, N) in shapes:
        say("bits", ls, rs, M, K, N)
        lhs = rand_values(rng, (M, K), 1, ls)
        rhs = rand_values(rng, (N, K), 1, rs)
        want = O.wrap_i32(O.blas_gemm_exact(O.values_of(lhs, 1, ls), O.values_of(rhs, 1, rs)))
        got = run_gemm(gbs, lhs, rhs, 1, 1, ls, rs, algo="mfma_bits")
        assert np.array_equal(got, want)
    for (M, K, N, ka) in [(50, 150, 70, 192), (77, 8190, 65, 8256), (140, 9000, 20, 9024)]:
        say("bits odd pitch", ls, rs, M, K, N, ka)
        lhs = rand_values(rng, (M, K), 1, ls)
        rhs = rand_values(rng, (N, K), 1, rs)
        got = run_gemm(gbs, lhs, rhs, 1, 1, ls, rs, algo="mfma_bits", geom=(M, N, M, ka, N, ka))
        assert np.array_equal(got, O.wrap_i32(O.naive_gemm_i64(lhs, rhs)))


def test_bits_engine_refuses_wider_operands(gbs, orc):
    rng = np.random.default_rng(1310)
    lhs = rand_values(rng, (8, 100), 2, False)
    rhs = rand_values(rng, (8, 100), 1, True)
    with pytest.raises(gbs.GbsError):
        run_gemm(gbs, lhs, rhs, 2, 1, False, True, algo="mfma_bits")


def test_bits_engine_mnist_golden(gbs, orc):
    """The reference's MNIST layer (test/mnistdata.h: bipolar weights x a
    2-bit unsigned input) through 1-bit products only: the input is its
    two bit-planes x = x0 + 2 x1, so W x0 + 2 W x1 must equal the
    reference's golden result."""
    w, x, golden = mnist()
    x = x.astype(np.int64).reshape(1, -1)
    say("bits mnist")
    y0 = run_gemm(gbs, w, (x & 1).astype(np.int8), 1, 1, True, False, algo="mfma_bits")
    y1 = run_gemm(gbs, w, (x >> 1).astype(np.int8), 1, 1, True, False, algo="mfma_bits")
    assert np.array_equal(y0.astype(np.int64) + 2 * y1.astype(np.int64),
                          golden.reshape(1, -1).astype(np.int64))


def test_auto_takes_bits_engine_for_shallow_1bit_products(gbs, orc):
    """AUTO's rule for the 1-bit planes engine: both sides 1-bit, depth at
    most 8192 and at most 1024 tiles of 64 x 64; the FP4 copies otherwise."""
    def sel(M, K, N, lb=1, rb=1):
        l, r = gbs.DeviceMatrix(lb, M, K, True), gbs.DeviceMatrix(rb, N, K, False)
        try:
            return gbs.select(l, r)
        finally:
            l.close(), r.close()
    say("bits auto rule")
    assert sel(4096, 2048, 1024) == "mfma_bits"   # 64 x 16 = 1024 tiles
    assert sel(4096, 8192, 512) == "mfma_bits"    # the binary layers
    assert sel(4096, 8193, 512) == "mfma_fp4"     # deeper
    assert sel(4096, 2048, 1025) == "mfma_fp4"    # 1088 tiles
    assert sel(4096, 2048, 1024, rb=2) != "mfma_bits"


# ------------------------------------------------------------ int8 MFMA engine
@pytest.mark.parametrize("lb,rb", list(itertools.product([1, 2, 3, 4], repeat=2)))
@pytest.mark.parametrize("ls,rs", list(itertools.product([False, True], repeat=2)))
def test_mfma_all_encodings(gbs, orc, lb, rb, ls, rs):
    rng = np.random.default_rng(7000 + lb * 100 + rb * 10 + 2 * ls + rs)
    for (M, K, N) in [(3, 6, 9), (17, 34, 1), (130, 700, 67), (64, 1030, 200),
                      (257, 3000, 129)]:
        say("mfma", lb, rb, ls, rs, M, K, N)
        lhs = rand_values(rng, (M, K), lb, ls)
        rhs = rand_values(rng, (N, K), rb, rs)
        got = run_gemm(gbs, lhs, rhs, lb, rb, ls, rs, algo="mfma")
        assert np.array_equal(got, expected(lhs, rhs, lb, rb, ls, rs))


def test_mfma_split_k_and_large(gbs, orc):
    rng = np.random.default_rng(77)
    for (M, K, N, lb, rb, ls, rs) in [(100, 20000, 40, 1, 1, True, True),
                                      (300, 8192, 520, 4, 4, True, False),
                                      (4096, 512, 256, 2, 2, False, True)]:
        lhs = rand_values(rng, (M, K), lb, ls)
        rhs = rand_values(rng, (N, K), rb, rs)
        got = run_gemm(gbs, lhs, rhs, lb, rb, ls, rs, algo="mfma")
        want = O.wrap_i32(O.blas_gemm_exact(O.values_of(lhs, lb, ls), O.values_of(rhs, rb, rs)))
        assert np.array_equal(got, want)


def test_mfma_sees_reimported_and_host_written_planes(gbs, orc):
    rng = np.random.default_rng(5)
    a = rand_values(rng, (70, 300), 2, True)
    b = rand_values(rng, (50, 300), 3, False)
    l, r = gbs.from_values(a, 2, True), gbs.from_values(b, 3, False)
    assert np.array_equal(gbs.gemm(l, r, "mfma"), expected(a, b, 2, 3, True, False))
    b2 = rand_values(rng, (50, 300), 3, False)
    r.import_regular(b2)
    assert np.array_equal(gbs.gemm(l, r, "mfma"), expected(a, b2, 2, 3, True, False))
    a2 = rand_values(rng, (70, 300), 2, True)
    l.set_planes(O.import_regular(a2, 2, True).planes)
    assert np.array_equal(gbs.gemm(l, r, "mfma"), expected(a2, b2, 2, 3, True, False))


def test_mfma_refuses_inexact(gbs):
    l = gbs.DeviceMatrix(8, 4, 64, False)   # unsigned 8-bit codes do not fit int8
    r = gbs.DeviceMatrix(2, 4, 64, False)
    with pytest.raises(gbs.GbsError, match="int8 matrix-core engine cannot hold"):
        gbs.gemm(l, r, "mfma")
    assert gbs.select(l, r) == "popcount"


def test_config_entry_point_refuses_bad_arguments(gbs):
    """The forced-shape entry point (the autotuner's) checks its arguments
    before any launch: null result, shape and split ranges, format fit."""
    n_classes, n_shapes = gbs.tune_classes()
    l = gbs.from_values(np.ones((20, 128), np.int8), 1, False)
    r = gbs.from_values(np.ones((10, 128), np.int8), 1, False)
    ws = gbs.Workspace()
    ws.reserve(l, r)
    L = gbs.lib()
    d = gbs.DeviceBuffer(4 * 20 * 10)
    assert L.gbs_gemm_async_config(ws.h, l.h, r.h, None, gbs.MFMA, 0, 1, None) == 1
    assert b"null result" in L.gbs_last_error()
    for shape, splits in [(-1, 1), (n_shapes, 1), (0, 0)]:
        assert L.gbs_gemm_async_config(ws.h, l.h, r.h, d.ptr, gbs.MFMA, shape, splits, None) == 1
        assert b"out of range" in L.gbs_last_error()
    assert L.gbs_gemm_async_config(ws.h, l.h, r.h, d.ptr, gbs.POPCOUNT, 0, 1, None) == 1
    assert b"matrix-core engine" in L.gbs_last_error()
    w = gbs.from_values(np.full((20, 128), 7, np.int8), 3, False)   # unsigned 3-bit: no FP4
    assert L.gbs_gemm_async_config(ws.h, w.h, r.h, d.ptr, gbs.MFMA_FP4, 0, 1, None) == 5
    gbs.check(L.gbs_gemm_async_config(ws.h, l.h, r.h, d.ptr, gbs.MFMA, 0, 1, None))
    gbs.sync()
    assert np.array_equal(d.to_host(np.int32, (10, 20)), np.full((10, 20), 128, np.int32))


def test_auto_matches_both_engines(gbs, orc):
    rng = np.random.default_rng(8)
    for (M, K, N, lb, rb, ls, rs) in [(512, 2048, 512, 1, 1, True, True),
                                      (256, 2304, 1024, 2, 3, True, False),
                                      (1000, 1000, 3, 4, 4, True, True)]:
        lhs = rand_values(rng, (M, K), lb, ls)
        rhs = rand_values(rng, (N, K), rb, rs)
        l, r = gbs.from_values(lhs, lb, ls), gbs.from_values(rhs, rb, rs)
        ra, rp, rm = gbs.gemm(l, r, "auto"), gbs.gemm(l, r, "popcount"), gbs.gemm(l, r, "mfma")
        assert np.array_equal(ra, rp) and np.array_equal(rp, rm)


def test_mfma_every_tile_shape_and_split(gbs, orc):
    """Each compiled tile shape, with and without a K split, on ragged
    shapes (the autotuner may pick any of them)."""
    n_classes, n_shapes = gbs.tune_classes()
    rng = np.random.default_rng(99)
    ws = gbs.Workspace()
    for (M, K, N, lb, rb, ls, rs) in [(300, 1100, 70, 3, 2, True, False),
                                      (129, 2048, 257, 1, 1, True, True),
                                      (64, 320, 16, 4, 4, True, True)]:
        lhs = rand_values(rng, (M, K), lb, ls)
        rhs = rand_values(rng, (N, K), rb, rs)
        l, r = gbs.from_values(lhs, lb, ls), gbs.from_values(rhs, rb, rs)
        ws.reserve(l, r)
        want = expected(lhs, rhs, lb, rb, ls, rs)
        d = gbs.DeviceBuffer(4 * M * N)
        for shape in range(n_shapes):
            for splits in (1, 3):
                say("shape", shape, "splits", splits, M, K, N)
                d.fill_zero()
                gbs.check(gbs.lib().gbs_gemm_async_config(ws.h, l.h, r.h, d.ptr, gbs.MFMA, shape,
                                                          splits,
                                                          None))
                gbs.sync()
                assert np.array_equal(d.to_host(np.int32, (N, M)), want), (shape, splits)


# ------------------------------------------------------------ FP4 MFMA engine
# operands an FP4 code holds exactly: bipolar, unsigned <= 2 bits, signed <= 3 bits
F4_ENC = [(1, True), (1, False), (2, False), (2, True), (3, True)]


@pytest.mark.parametrize("lenc", F4_ENC)
@pytest.mark.parametrize("renc", F4_ENC)
def test_fp4_all_encodings(gbs, orc, lenc, renc):
    (lb, ls), (rb, rs) = lenc, renc
    rng = np.random.default_rng(9000 + lb * 100 + rb * 10 + 2 * ls + rs)
    for (M, K, N) in [(3, 6, 9), (17, 34, 1), (130, 700, 67), (64, 1030, 200),
                      (257, 3000, 129)]:
        say("fp4", lb, rb, ls, rs, M, K, N)
        lhs = rand_values(rng, (M, K), lb, ls)
        rhs = rand_values(rng, (N, K), rb, rs)
        got = run_gemm(gbs, lhs, rhs, lb, rb, ls, rs, algo="mfma_fp4")
        assert np.array_equal(got, expected(lhs, rhs, lb, rb, ls, rs))


def test_fp4_refuses_what_it_cannot_hold(gbs):
    for (lb, ls) in [(3, False), (4, True)]:
        l = gbs.DeviceMatrix(lb, 4, 64, ls)
        r = gbs.DeviceMatrix(1, 4, 64, True)
        with pytest.raises(gbs.GbsError, match="FP4 matrix-core engine cannot hold"):
            gbs.gemm(l, r, "mfma_fp4")
    # K * 4 * 4 reaches 2^24: fp32 sums could round
    K = 1 << 20
    l = gbs.DeviceMatrix(3, 2, K, True)
    r = gbs.DeviceMatrix(3, 2, K, True)
    with pytest.raises(gbs.GbsError, match="FP4"):
        gbs.gemm(l, r, "mfma_fp4")
    assert gbs.select(l, r) != "mfma_fp4"


def test_fp4_exact_just_below_the_bound(gbs, orc):
    """Sums of magnitude close to 2^24 with odd low bits: any fp32 rounding
    inside the instruction or across the K loop would show."""
    rng = np.random.default_rng(2024)
    K = (1 << 20) - 64                      # K * 4 * 4 < 2^24
    M, N = 48, 20
    lhs = rng.choice(np.array([-4, 3, -3, 1], np.int8), size=(M, K), p=[0.85, 0.05, 0.05, 0.05])
    rhs = rng.choice(np.array([-4, 3, -1], np.int8), size=(N, K), p=[0.9, 0.05, 0.05])
    l, r = gbs.from_values(lhs, 3, True), gbs.from_values(rhs, 3, True)
    assert gbs.select(l, r) == "mfma_fp4" or gbs.select(l, r) == "mfma"
    got = gbs.gemm(l, r, "mfma_fp4")
    want = O.wrap_i32(O.blas_gemm_exact(lhs.astype(np.int64), rhs.astype(np.int64)))
    assert np.abs(want).max() > (1 << 23)
    assert np.array_equal(got, want)
    # bipolar, deep: every sum is K - 2 * (disagreements)
    K2 = 300001
    a = rng.integers(0, 2, size=(33, K2)).astype(np.int8) * 2 - 1
    b = np.where(rng.random((7, K2)) < 0.97, 1, -1).astype(np.int8)
    la, lb_ = gbs.from_values(a, 1, True), gbs.from_values(b, 1, True)
    want2 = O.wrap_i32(O.blas_gemm_exact(a.astype(np.int64), b.astype(np.int64)))
    assert np.array_equal(gbs.gemm(la, lb_, "mfma_fp4"), want2)


def test_fp4_sees_reimported_and_host_written_planes(gbs, orc):
    rng = np.random.default_rng(15)
    a = rand_values(rng, (70, 300), 2, True)
    b = rand_values(rng, (50, 300), 2, False)
    l, r = gbs.from_values(a, 2, True), gbs.from_values(b, 2, False)
    assert np.array_equal(gbs.gemm(l, r, "mfma_fp4"), expected(a, b, 2, 2, True, False))
    b2 = rand_values(rng, (50, 300), 2, False)
    r.import_regular(b2)
    assert np.array_equal(gbs.gemm(l, r, "mfma_fp4"), expected(a, b2, 2, 2, True, False))
    a2 = rand_values(rng, (70, 300), 2, True)
    l.set_planes(O.import_regular(a2, 2, True).planes)
    assert np.array_equal(gbs.gemm(l, r, "mfma_fp4"), expected(a2, b2, 2, 2, True, False))
    # the int8 copy of the same matrices stays coherent too
    assert np.array_equal(gbs.gemm(l, r, "mfma"), expected(a2, b2, 2, 2, True, False))


def test_fp4_every_tile_shape_and_split(gbs, orc):
    n_classes, n_shapes = gbs.tune_classes()
    rng = np.random.default_rng(199)
    ws = gbs.Workspace()
    for (M, K, N, lb, rb, ls, rs) in [(300, 1100, 70, 3, 2, True, False),
                                      (129, 2048, 257, 1, 1, True, True),
                                      (64, 320, 16, 1, 2, False, True)]:
        lhs = rand_values(rng, (M, K), lb, ls)
        rhs = rand_values(rng, (N, K), rb, rs)
        l, r = gbs.from_values(lhs, lb, ls), gbs.from_values(rhs, rb, rs)
        ws.reserve(l, r)
        want = expected(lhs, rhs, lb, rb, ls, rs)
        d = gbs.DeviceBuffer(4 * M * N)
        for shape in range(n_shapes):
            for splits in (1, 3):
                say("fp4 shape", shape, "splits", splits, M, K, N)
                d.fill_zero()
                gbs.check(gbs.lib().gbs_gemm_async_config(ws.h, l.h, r.h, d.ptr, gbs.MFMA_FP4,
                                                          shape, splits, None))
                gbs.sync()
                assert np.array_equal(d.to_host(np.int32, (N, M)), want), (shape, splits)


# ------------------------------------------------------------ FP6 MFMA engine
# operands an FP6 code holds exactly: bipolar, unsigned <= 4 bits, signed <= 5 bits
F6_ENC = [(1, True), (2, False), (3, False), (3, True), (4, False), (4, True), (5, True)]


@pytest.mark.parametrize("lenc", F6_ENC)
@pytest.mark.parametrize("renc", F6_ENC)
def test_fp6_all_encodings(gbs, orc, lenc, renc):
    (lb, ls), (rb, rs) = lenc, renc
    rng = np.random.default_rng(11000 + lb * 100 + rb * 10 + 2 * ls + rs)
    for (M, K, N) in [(3, 6, 9), (17, 34, 1), (130, 700, 67), (64, 1030, 200),
                      (257, 3000, 129)]:
        say("fp6", lb, rb, ls, rs, M, K, N)
        lhs = rand_values(rng, (M, K), lb, ls)
        rhs = rand_values(rng, (N, K), rb, rs)
        got = run_gemm(gbs, lhs, rhs, lb, rb, ls, rs, algo="mfma_fp6")
        assert np.array_equal(got, expected(lhs, rhs, lb, rb, ls, rs))


def test_fp6_refuses_what_it_cannot_hold(gbs):
    for (lb, ls) in [(5, False), (6, True)]:
        l = gbs.DeviceMatrix(lb, 4, 64, ls)
        r = gbs.DeviceMatrix(1, 4, 64, True)
        with pytest.raises(gbs.GbsError, match="FP6 matrix-core engine cannot hold"):
            gbs.gemm(l, r, "mfma_fp6")
    # K * 8 * 15 reaches 2^24
    K = 140000
    l = gbs.DeviceMatrix(4, 2, K, True)
    r = gbs.DeviceMatrix(4, 2, K, False)
    with pytest.raises(gbs.GbsError, match="FP6"):
        gbs.gemm(l, r, "mfma_fp6")
    assert gbs.select(l, r) != "mfma_fp6"


def test_fp6_exact_just_below_the_bound(gbs, orc):
    rng = np.random.default_rng(2025)
    K = 139000                              # K * 8 * 15 < 2^24
    M, N = 40, 24
    lhs = rng.choice(np.array([-8, 7, -5, 3], np.int8), size=(M, K), p=[0.9, 0.04, 0.03, 0.03])
    rhs = rng.choice(np.array([15, 13, 1, 0], np.int8), size=(N, K), p=[0.9, 0.04, 0.03, 0.03])
    l, r = gbs.from_values(lhs, 4, True), gbs.from_values(rhs, 4, False)
    got = gbs.gemm(l, r, "mfma_fp6")
    want = O.wrap_i32(O.blas_gemm_exact(lhs.astype(np.int64), rhs.astype(np.int64)))
    assert np.abs(want).max() > (1 << 23)
    assert np.array_equal(got, want)
    # signed 5-bit reaches -16, the largest FP6 magnitude used
    K2 = 60000                              # K * 16 * 16 < 2^24
    a = rng.choice(np.array([-16, 15, -3], np.int8), size=(20, K2), p=[0.9, 0.05, 0.05])
    b = rng.choice(np.array([-16, 11, 2], np.int8), size=(12, K2), p=[0.9, 0.05, 0.05])
    la, lb_ = gbs.from_values(a, 5, True), gbs.from_values(b, 5, True)
    want2 = O.wrap_i32(O.blas_gemm_exact(a.astype(np.int64), b.astype(np.int64)))
    assert np.abs(want2).max() > (1 << 23)
    assert np.array_equal(gbs.gemm(la, lb_, "mfma_fp6"), want2)


def test_fp6_every_tile_shape_and_split(gbs, orc):
    n_classes, n_shapes = gbs.tune_classes()
    rng = np.random.default_rng(299)
    ws = gbs.Workspace()
    for (M, K, N, lb, rb, ls, rs) in [(300, 1100, 70, 3, 4, False, True),
                                      (129, 2048, 257, 4, 4, True, False),
                                      (64, 320, 16, 5, 1, True, True)]:
        lhs = rand_values(rng, (M, K), lb, ls)
        rhs = rand_values(rng, (N, K), rb, rs)
        l, r = gbs.from_values(lhs, lb, ls), gbs.from_values(rhs, rb, rs)
        ws.reserve(l, r)
        want = expected(lhs, rhs, lb, rb, ls, rs)
        d = gbs.DeviceBuffer(4 * M * N)
        for shape in range(n_shapes):
            for splits in (1, 3):
                say("fp6 shape", shape, "splits", splits, M, K, N)
                d.fill_zero()
                gbs.check(gbs.lib().gbs_gemm_async_config(ws.h, l.h, r.h, d.ptr, gbs.MFMA_FP6,
                                                          shape, splits, None))
                gbs.sync()
                assert np.array_equal(d.to_host(np.int32, (N, M)), want), (shape, splits)


def test_auto_picks_format_from_support_and_measured_times(gbs, tmp_path):
    import os
    # 1-bit x 1-bit at this depth and tile count: the matrix cores on the
    # planes themselves (no copies)
    cases = [((1, True), (1, True), "mfma_bits"), ((2, False), (3, True), "mfma_fp4"),
             ((3, False), (2, True), "mfma"), ((4, True), (4, False), "mfma"),
             ((6, True), (2, False), "mfma")]
    M, N, K = 2048, 512, 2048
    mats = {}

    def mat(b, sg, rows):
        key = (b, sg, rows)
        if key not in mats:
            mats[key] = gbs.DeviceMatrix(b, rows, K, sg)
        return mats[key]

    table = os.path.join(os.path.dirname(gbs.__file__), "tune_table.txt")
    cls = gbs.tune_class_of(M, N, K)
    try:
        gbs.tune_set(None)              # no measured times: FP6 is never preferred
        for (lb, ls), (rb, rs), want in cases:
            assert gbs.select(mat(lb, ls, M), mat(rb, rs, N)) == want, ((lb, ls), (rb, rs))
        # a table that measured FP6 faster than int8 for this class
        t = tmp_path / "table.txt"
        t.write_text(f"{cls} 4 1 0 10.0\n{cls} 4 1 2 9.0\n")
        gbs.tune_load(str(t))
        assert gbs.select(mat(3, False, M), mat(2, True, N)) == "mfma_fp6"
        assert gbs.select(mat(4, True, M), mat(4, False, N)) == "mfma_fp6"
        assert gbs.select(mat(6, True, M), mat(2, False, N)) == "mfma"      # FP6 cannot hold it
        assert gbs.select(mat(1, True, M), mat(1, True, N)) == "mfma_bits"
        t.write_text(f"{cls} 4 1 0 10.0\n{cls} 4 1 2 9.8\n")             # within 5 %: int8
        gbs.tune_load(str(t))
        assert gbs.select(mat(3, False, M), mat(2, True, N)) == "mfma"
    finally:
        gbs.tune_load(table)


def test_random_sweep_every_engine(gbs, orc):
    """A seeded sweep of random shapes (log-uniform 1..1500 rows, 1..5000
    columns) and encodings (1..8 bits, either signedness) through AUTO and
    every engine that holds the product exactly, each against the exact
    product of the values the oracle's import denotes (float64 BLAS, exact
    at these magnitudes; no int32 wrap is reachable)."""
    rng = np.random.default_rng(20261203)
    engines = ["popcount", "mfma", "mfma_fp4", "mfma_fp6", "mfma_bits"]
    ran = {e: 0 for e in engines}
    for case in range(40):
        M, N = (int(np.exp(rng.uniform(0, np.log(1500)))) for _ in range(2))
        K = int(np.exp(rng.uniform(0, np.log(5000))))
        lb, rb = (int(rng.integers(1, 9)) for _ in range(2))
        ls, rs = (bool(rng.integers(0, 2)) for _ in range(2))
        lhs = rand_values(rng, (M, K), lb, ls, np.int16)
        rhs = rand_values(rng, (N, K), rb, rs, np.int16)
        l, r = gbs.from_values(lhs, lb, ls), gbs.from_values(rhs, rb, rs)
        want = O.wrap_i32(O.blas_gemm_exact(O.values_of(lhs, lb, ls), O.values_of(rhs, rb, rs)))
        say("sweep", case, M, K, N, lb, rb, ls, rs, gbs.select(l, r))
        assert np.array_equal(gbs.gemm(l, r, "auto"), want), (case, "auto")
        for e in engines:
            try:
                got = gbs.gemm(l, r, e)
            except gbs.GbsError as err:
                assert e != "popcount" and err.code == 5, (case, e, err)  # cannot hold it exactly
                continue
            ran[e] += 1
            assert np.array_equal(got, want), (case, e)
        l.close(), r.close()
    # the 1-bit engine holds a product only when both sides are 1-bit
    assert min(v for e, v in ran.items() if e != "mfma_bits") >= 3, ran
