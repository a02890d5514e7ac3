"""The planes matrix-core engine (GBS_ALGO_MFMA_PLANES, gbs_planes.hip):
multi-bit operands straight from their bit-planes, one side in 2-bit chunks
as FP4 codes, the other as FP4 codes of its whole values.  Every result is
compared bit for bit with the oracle's gemm (oracle/gbs_oracle.c, pinned to
the compiled reference in tests/test_oracle.py) on the oracle's own packing
of the same values: every plane-side width 1..4 (unsigned, signed,
bipolar) against every code-side encoding that fits an FP4 code (1-bit,
bipolar, 2-bit unsigned, 2- and 3-bit signed) and every one taken as two
2-bit chunks (3- and 4-bit unsigned, 4-bit signed), in both orientations,
at ragged shapes, depths with partly filled super-steps and odd row
pitches, each of the compiled tile shapes, and sums at 99 % of the 2^24
bound."""
import itertools

import numpy as np
import pytest

from conftest import rand_values
from oracle import orc as O

pytestmark = pytest.mark.gpu

PLANE_ENC = [(b, s) for b in (1, 2, 3, 4) for s in (False, True)]   # (1, True) = bipolar
CODE_ENC = [(1, False), (1, True), (2, False), (2, True), (3, True)]
CODE_ENC2 = [(3, False), (4, False), (4, True)]   # two 2-bit chunks


def _check(gbs, a, b, lb, ls, rb, rs, ncols_a=None):
    l = gbs.from_values(a, lb, ls, ncols_a=ncols_a)
    r = gbs.from_values(b, rb, rs, ncols_a=ncols_a)
    want = O.gemm(O.import_regular(a, lb, ls, ncols_a=l.ncols_a),
                  O.import_regular(b, rb, rs, ncols_a=r.ncols_a))
    got = gbs.gemm(l, r, "mfma_planes")
    l.close()
    r.close()
    return got, want


@pytest.mark.parametrize("plane,code", list(itertools.product(PLANE_ENC, CODE_ENC)))
@pytest.mark.parametrize("code_is_lhs", [True, False])
def test_planes_every_encoding_pair(gbs, orc, plane, code, code_is_lhs):
    pb, ps = plane
    cb, cs = code
    if pb == 1 and ps and cb == 1 and cs:
        pytest.skip("bipolar x bipolar is the 1-bit engine's product")
    rng = np.random.default_rng(pb * 100 + ps * 10 + cb * 3 + cs + 1000 * code_is_lhs)
    # the code side has fewer rows, so the engine keeps it as the code side
    for NP, NC, K in [(300, 70, 2304), (129, 17, 1000), (64, 130, 64)]:
        if NC > NP:
            NP, NC = NC + 40, NC
        P = rand_values(rng, (NP, K), pb, ps)
        Cv = rand_values(rng, (NC, K), cb, cs)
        if code_is_lhs:
            got, want = _check(gbs, Cv, P, cb, cs, pb, ps)
        else:
            got, want = _check(gbs, P, Cv, pb, ps, cb, cs)
        assert np.array_equal(got, want), (plane, code, code_is_lhs, NP, NC, K)


@pytest.mark.parametrize("K", [1, 63, 65, 1023, 1024, 1025, 2047, 3000, 4160])
def test_planes_depths_and_tails(gbs, orc, K):
    """Depths across word, 128-bit and super-step (1024) boundaries."""
    rng = np.random.default_rng(K)
    a = rand_values(rng, (200, K), 2, True)      # code side (lhs, fewer rows)
    b = rand_values(rng, (333, K), 3, False)     # plane side
    got, want = _check(gbs, a, b, 2, True, 3, False)
    assert np.array_equal(got, want)


def test_planes_odd_row_pitch(gbs, orc):
    """Rows of an odd number of 64-bit words (8-byte aligned only)."""
    rng = np.random.default_rng(5)
    for K, ncols_a in [(192, 192), (320, 448), (1000, 1088)]:
        a = rand_values(rng, (50, K), 2, False)
        b = rand_values(rng, (260, K), 4, True)
        got, want = _check(gbs, a, b, 2, False, 4, True, ncols_a=ncols_a)
        assert np.array_equal(got, want), (K, ncols_a)


@pytest.mark.parametrize("NC", [5, 16, 33, 64, 65, 128, 129, 300])
@pytest.mark.parametrize("pb", [2, 4])
def test_planes_tile_shapes(gbs, orc, NC, pb):
    """Code-side row counts that pick each compiled tile (and partly filled
    ones), with enough plane-side rows for the widest 3..4-plane tile."""
    rng = np.random.default_rng(NC * 7 + pb)
    NP = 8300 if pb == 4 and NC > 32 else 700
    a = rand_values(rng, (NC, 384), 3, True)
    b = rand_values(rng, (NP, 384), pb, False)
    got, want = _check(gbs, a, b, 3, True, pb, False)
    assert np.array_equal(got, want)


def test_planes_sums_near_the_bound(gbs, orc):
    """|sum| = K * 15 * 4 at 99 % of 2^24: plane side all 15 (4-bit
    unsigned), code side all -4 (3-bit signed), odd values mixed in."""
    K = 276480
    assert K * 15 * 4 < (1 << 24) and K * 15 * 4 > 0.98 * (1 << 24)
    a = np.full((20, K), -4, np.int8)
    a[:, ::997] = 3
    b = np.full((40, K), 15, np.int8)
    b[3, ::5] = 13
    got, want = _check(gbs, a, b, 3, True, 4, False)
    assert np.array_equal(got, want)
    assert abs(int(want.astype(np.int64).min())) > 0.97 * (1 << 24)


@pytest.mark.parametrize("plane,code", list(itertools.product(PLANE_ENC, CODE_ENC2)))
@pytest.mark.parametrize("code_is_lhs", [True, False])
def test_planes_two_chunk_code_side(gbs, orc, plane, code, code_is_lhs):
    """Operands neither of which fits one FP4 code: the one with fewer rows
    is taken as two 2-bit chunks (4 instructions per k-step against a
    two-chunk plane side), the other streamed as planes."""
    pb, ps = plane
    cb, cs = code
    if pb <= 2 or (pb == 3 and ps):
        # the plane operand fits one FP4 code: the engine makes it the code side
        pb, ps = (4, ps) if pb <= 2 else (4, True)
    rng = np.random.default_rng(pb * 100 + ps * 10 + cb * 3 + cs + 2000 * code_is_lhs)
    for NP, NC, K in [(300, 70, 2304), (129, 17, 1000), (200, 64, 64), (96, 80, 3000)]:
        P = rand_values(rng, (NP, K), pb, ps)
        Cv = rand_values(rng, (NC, K), cb, cs)
        if code_is_lhs:
            got, want = _check(gbs, Cv, P, cb, cs, pb, ps)
        else:
            got, want = _check(gbs, P, Cv, pb, ps, cb, cs)
        assert np.array_equal(got, want), (plane, code, code_is_lhs, NP, NC, K)


@pytest.mark.parametrize("K", [1, 65, 1024, 1025, 2047, 4160])
def test_planes_two_chunk_depths(gbs, orc, K):
    """4-bit signed weights x 4-bit unsigned activations (the conv product's
    4 x 4 form) across depth boundaries and an odd row pitch."""
    rng = np.random.default_rng(K + 17)
    a = rand_values(rng, (70, K), 4, True)
    b = rand_values(rng, (333, K), 4, False)
    got, want = _check(gbs, a, b, 4, True, 4, False)
    assert np.array_equal(got, want), K
    if K == 1025:
        got, want = _check(gbs, a, b, 4, True, 4, False, ncols_a=1088)
        assert np.array_equal(got, want)


def test_planes_two_chunk_sums_near_the_bound(gbs, orc):
    """Code side 4-bit unsigned (chunk maxima 3 + 4 x 3 = 15) against plane
    side 4-bit unsigned all 15: |sum| = K * 225 at 99 % of 2^24."""
    K = 73600
    assert K * 15 * 15 < (1 << 24) and K * 15 * 15 > 0.98 * (1 << 24)
    a = np.full((20, K), 15, np.int8)
    a[:, ::991] = 13
    b = np.full((40, K), 15, np.int8)
    b[5, ::7] = 14
    got, want = _check(gbs, a, b, 4, False, 4, False)
    assert np.array_equal(got, want)
    assert int(want.astype(np.int64).max()) > 0.97 * (1 << 24)


def test_planes_refuses_what_it_cannot_hold(gbs, orc):
    rng = np.random.default_rng(2)
    l = gbs.from_values(rand_values(rng, (32, 64), 5, True), 5, True)    # neither side fits
    r = gbs.from_values(rand_values(rng, (32, 64), 5, False), 5, False)
    with pytest.raises(gbs.GbsError) as e:
        gbs.gemm(l, r, "mfma_planes")
    assert e.value.code == 5
    big = gbs.from_values(np.full((16, 300000), 15, np.int8), 4, False)   # past 2^24
    cod = gbs.from_values(np.full((16, 300000), -4, np.int8), 3, True)
    with pytest.raises(gbs.GbsError):
        gbs.gemm(cod, big, "mfma_planes")
    for x in (l, r, big, cod):
        x.close()


def test_auto_reads_planes_when_copies_are_stale(gbs, orc):
    """AUTO on a large product whose value copies would have to be rebuilt
    (freshly imported activations) takes the planes engine; once a copy
    engine has made the copies it takes them; both bit-exact."""
    from oracle import orc as O
    rng = np.random.default_rng(77)
    M, K, N = 256, 2304, 8192   # 2 * M * K * N = 9.7e9 ops
    a = rng.integers(-2, 2, (M, K)).astype(np.int8)
    b = rng.integers(0, 8, (N, K)).astype(np.int8)
    want = O.wrap_i32(O.blas_gemm_exact(O.values_of(a, 2, True), O.values_of(b, 3, False)))
    l, r = gbs.from_values(a, 2, True), gbs.from_values(b, 3, False)
    assert gbs.select(l, r) == "mfma_planes"
    assert np.array_equal(gbs.gemm(l, r, "auto"), want)
    assert np.array_equal(gbs.gemm(l, r, "mfma"), want)       # builds the int8 copies
    assert gbs.select(l, r) == "mfma"
    assert np.array_equal(gbs.gemm(l, r, "auto"), want)
    b2 = rng.integers(0, 8, (N, K)).astype(np.int8)
    r.import_regular(b2)                                      # copies stale again
    assert gbs.select(l, r) == "mfma_planes"
    want2 = O.wrap_i32(O.blas_gemm_exact(O.values_of(a, 2, True), O.values_of(b2, 3, False)))
    assert np.array_equal(gbs.gemm(l, r, "auto"), want2)
    # small products keep the copy engines whatever the copies' state
    ls, rs = gbs.from_values(a[:, :512], 2, True), gbs.from_values(b[:256, :512], 3, False)
    assert gbs.select(ls, rs) != "mfma_planes"
    for m in (l, r, ls, rs):
        m.close()
