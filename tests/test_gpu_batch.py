"""Batched persistent launch: every product equals its own single-product
result and the oracle, per-call and graph-replayed runs agree."""
import numpy as np
import pytest

from gemmbitserial_amd import workloads as W
from oracle import orc as O

pytestmark = pytest.mark.gpu


def build(gbs, probs):
    items = []
    for p in probs:
        l = gbs.from_values(p["lhs"], p["lb"], p["ls"])
        r = gbs.from_values(p["rhs"], p["rb"], p["rs"])
        d = gbs.DeviceBuffer(4 * max(1, p["M"] * p["N"]))
        items.append((l, r, d))
    return items


def exact(p):
    return O.wrap_i32(O.blas_gemm_exact(O.values_of(p["lhs"], p["lb"], p["ls"]),
                                        O.values_of(p["rhs"], p["rb"], p["rs"])))


def results(items, probs):
    return [d.to_host(np.int32, (p["N"], p["M"])) for (l, r, d), p in zip(items, probs)]


def zero(items):
    for _, _, d in items:
        d.fill_zero()


@pytest.fixture(params=[2, 1, 0], ids=["planes_all", "planes_skinny", "copies"])
def planes(request, gbs):
    """Runs a test with every product the planes group holds in it, with
    skinny products beyond FP4 in it, and with every product from value
    copies (the default)."""
    gbs.batch_set_planes(request.param)
    yield request.param
    gbs.batch_set_planes(gbs.PLANES_NONE)


def test_batch_mixed_small(gbs, orc, planes):
    probs = W.batched_uneven(count=48, seed=11, scale=0.25)
    # edge cases: an 8-bit unsigned product (popcount fallback), K = 0, N = 0
    rng = np.random.default_rng(3)
    probs.append(dict(M=33, K=200, N=7, lb=8, rb=2, ls=False, rs=True,
                      lhs=rng.integers(0, 256, (33, 200)).astype(np.uint8),
                      rhs=W.values(rng, (7, 200), 2, True)))
    probs.append(dict(M=5, K=0, N=3, lb=2, rb=2, ls=False, rs=False,
                      lhs=np.zeros((5, 0), np.int8), rhs=np.zeros((3, 0), np.int8)))
    probs.append(dict(M=9, K=64, N=0, lb=1, rb=1, ls=True, rs=True,
                      lhs=W.values(rng, (9, 64), 1, True), rhs=np.zeros((0, 64), np.int8)))
    items = []
    for p in probs:
        l = gbs.DeviceMatrix(p["lb"], p["M"], p["K"], p["ls"])
        r = gbs.DeviceMatrix(p["rb"], p["N"], p["K"], p["rs"])
        if p["M"] * p["K"]:
            l.import_regular(p["lhs"])
        if p["N"] * p["K"]:
            r.import_regular(p["rhs"])
        items.append((l, r, gbs.DeviceBuffer(4 * max(1, p["M"] * p["N"]))))
    b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    info = b.info()
    assert info["separate"] >= 2 and info["persistent"] >= 40
    want = [exact(p) if p["K"] else np.zeros((p["N"], p["M"]), np.int32) for p in probs]
    b.run()
    gbs.sync()
    for g, w in zip(results(items, probs), want):
        assert np.array_equal(g, w)
    zero(items)
    b.run_per_call()
    gbs.sync()
    for g, w in zip(results(items, probs), want):
        assert np.array_equal(g, w)
    # graph: captured once, replayed with eager work and syncs in between
    for rep in range(3):
        zero(items)
        gbs.sync()
        b.run_graph()
        gbs.sync()
        for g, w in zip(results(items, probs), want):
            assert np.array_equal(g, w), f"graph replay {rep}"
    b.close()


def test_batch_order_does_not_matter(gbs, orc, planes):
    probs = W.batched_uneven(count=24, seed=5, scale=0.2)
    items = build(gbs, probs)
    perm = np.random.default_rng(0).permutation(len(items))
    b1 = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    b1.run()
    gbs.sync()
    first = results(items, probs)
    zero(items)
    b2 = gbs.Batch([(items[i][0], items[i][1], items[i][2].ptr) for i in perm])
    b2.run()
    gbs.sync()
    for g, f, p in zip(results(items, probs), first, probs):
        assert np.array_equal(g, f)
        assert np.array_equal(g, exact(p))


def test_batch_full_size_c3(gbs, orc, planes):
    """The 512-product uneven workload at full size: every result exact."""
    probs = W.batched_uneven()
    items = build(gbs, probs)
    b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    groups = {b.plan_of(i)[0] for i in range(len(probs))}
    assert (gbs.BATCH_PLANES in groups) == (planes != gbs.PLANES_NONE)
    b.run_graph()
    gbs.sync()
    for g, p in zip(results(items, probs), probs):
        assert np.array_equal(g, exact(p))


def test_batch_caffenet_list(gbs, orc, planes):
    """The reference benchmark's CaffeNet GEMM list (2u x 2u): batched, graph
    replayed and per-call runs all equal the oracle's bit-serial product
    (which follows the reference's gemmBitSerial)."""
    probs = W.caffenet()
    items = build(gbs, probs)
    b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    want = [O.gemm(O.import_regular(p["lhs"], p["lb"], p["ls"]),
                   O.import_regular(p["rhs"], p["rb"], p["rs"])) for p in probs]
    for run in (b.run, b.run_graph, b.run_graph, b.run_per_call):
        zero(items)
        run()
        gbs.sync()
        for g, w in zip(results(items, probs), want):
            assert np.array_equal(g, w)


def test_batch_sees_reimported_operands(gbs, orc, planes):
    """An operand re-imported after the plan (same shape, new values): plain
    and graph-replayed runs both use the new values, in every copy format."""
    rng = np.random.default_rng(21)
    shapes = [(96, 640, 40, 1, 1, True, True),     # FP4 copies
              (70, 512, 24, 4, 3, True, False),    # FP6 copies
              (50, 384, 90, 6, 2, True, True)]     # int8 copies
    probs = []
    for (M, K, N, lb, rb, ls, rs) in shapes:
        probs.append(dict(M=M, K=K, N=N, lb=lb, rb=rb, ls=ls, rs=rs,
                          lhs=W.values(rng, (M, K), lb, ls), rhs=W.values(rng, (N, K), rb, rs)))
    items = build(gbs, probs)
    b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    for run in (b.run, b.run_graph):
        for step in range(2):
            for (l, r, _), p in zip(items, probs):
                p["lhs"] = W.values(rng, (p["M"], p["K"]), p["lb"], p["ls"])
                l.import_regular(p["lhs"])
                if step:
                    p["rhs"] = W.values(rng, (p["N"], p["K"]), p["rb"], p["rs"])
                    r.import_regular(p["rhs"])
            zero(items)
            gbs.sync()
            run()
            gbs.sync()
            for g, p in zip(results(items, probs), probs):
                assert np.array_equal(g, exact(p)), (run.__name__, step, p["lb"], p["rb"])
    b.close()


def test_batch_plans_follow_the_tile_table(gbs, orc):
    """Products of a class with a table entry are planned with that entry's
    tile shape (here forced to each carried shape in turn, with a fine and a
    coarse K piece) and stay bit-exact."""
    rng = np.random.default_rng(21)
    probs = [dict(M=300, K=2560, N=12, lb=2, rb=1, ls=False, rs=False),   # FP4, N <= 16
             dict(M=260, K=5120, N=100, lb=4, rb=3, ls=True, rs=False)]   # FP6, N <= 128
    for p in probs:
        p["lhs"] = W.values(rng, (p["M"], p["K"]), p["lb"], p["ls"])
        p["rhs"] = W.values(rng, (p["N"], p["K"]), p["rb"], p["rs"])
    items = build(gbs, probs)
    want = [exact(p) for p in probs]
    n = gbs.batch_tune_classes()
    saved = [[gbs.batch_tune_get(f, c) for c in range(n)] for f in range(3)]
    gbs.batch_set_planes(gbs.PLANES_NONE)  # the table governs the copy formats' launches
    try:
        for fmt, p, k in ((1, probs[0], 0), (2, probs[1], 1)):
            cls = gbs.batch_tune_class_of(p["N"], p["K"])
            for shape in gbs.batch_tune_shapes(fmt):
                for piece in (4, 32):
                    gbs.batch_tune_set(fmt, cls, shape, piece)
                    zero(items)
                    b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
                    f, s, splits, pc = b.plan_of(k)
                    assert (f, s) == (fmt, shape)
                    units = p["K"] // 2 // 256  # nominal 256-byte units of FP4 / FP6 copies
                    assert splits * pc >= units and (splits - 1) * pc < units
                    b.run()
                    gbs.sync()
                    for got, w in zip(results(items, probs), want):
                        assert np.array_equal(got, w), (fmt, shape, piece)
                    b.close()
    finally:
        gbs.batch_set_planes(gbs.PLANES_NONE)
        for f in range(3):
            for c in range(n):
                s, pc = saved[f][c]
                gbs.batch_tune_set(f, c, s, pc if s >= 0 else 16)
        for x in items:
            for y in x:
                y.close()


# (bits, signed) of every operand encoding the planes group takes on each
# side; 1-bit signed is bipolar
ENCODINGS = [(1, False), (1, True), (2, False), (2, True), (3, False), (3, True),
             (4, False), (4, True)]


def planes_form(lb, ls, rb, rs, N):
    """The tile form the planes group runs: plane-side chunks - 1, + 2 for a
    two-chunk code side (3-4 bits unsigned, 4 bits signed), + 4 for 64-row
    code-side tiles (more than 32 rhs rows)."""
    cc = 2 if (not rs and rb >= 3) or (rs and rb == 4) else 1
    return (lb + 1) // 2 - 1 + 2 * (cc - 1) + (4 if N > 32 else 0)


def test_batch_planes_group_every_form(gbs, orc):
    """Every lhs x rhs encoding pair the planes group takes (bipolar x
    bipolar stays with the FP4 copies), with narrow and wide rhs operands,
    ragged rows, depths that end mid-word, mid-super-step and on a
    super-step, and deep products cut into K pieces: each plan names the
    planes group and its tile form, and plain, graph-replayed and per-call
    runs equal the exact product."""
    rng = np.random.default_rng(41)
    depths = [64, 1000, 1024, 2113, 4096, 20000]
    probs = []
    for i, (lb, ls) in enumerate(ENCODINGS):
        for j, (rb, rs) in enumerate(ENCODINGS):
            M = int(rng.integers(20, 300))
            N = int(rng.choice([1, 7, 32, 33, 100]))
            K = depths[(i * 8 + j) % len(depths)]
            probs.append(dict(M=M, K=K, N=N, lb=lb, rb=rb, ls=ls, rs=rs,
                              lhs=W.values(rng, (M, K), lb, ls),
                              rhs=W.values(rng, (N, K), rb, rs)))
    items = build(gbs, probs)
    gbs.batch_set_planes(gbs.PLANES_ALL)
    try:
        b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    finally:
        gbs.batch_set_planes(gbs.PLANES_NONE)
    for k, p in enumerate(probs):
        g, form, splits, piece = b.plan_of(k)
        if p["lb"] == p["rb"] == 1 and p["ls"] and p["rs"]:
            assert g == 1, "bipolar x bipolar runs from FP4 copies"
            continue
        assert g == gbs.BATCH_PLANES, (p["lb"], p["ls"], p["rb"], p["rs"])
        assert form == planes_form(p["lb"], p["ls"], p["rb"], p["rs"], p["N"])
        # pieces of `piece` whole super-steps (1024 values) cover them once;
        # the last piece also takes the partly filled one
        full = (p["K"] + 63) // 64 // 16
        if full:
            assert (splits - 1) * piece < full <= splits * piece
        else:
            assert splits == 1
        if p["K"] == 20000:
            assert splits > 1
    want = [exact(p) for p in probs]
    for run in (b.run, b.run_graph, b.run_graph, b.run_per_call):
        zero(items)
        gbs.sync()
        run()
        gbs.sync()
        for g, w, p in zip(results(items, probs), want, probs):
            assert np.array_equal(g, w), (run.__name__, p["lb"], p["ls"], p["rb"], p["rs"],
                                          p["M"], p["K"], p["N"])
    b.close()


def test_batch_planes_group_sums_near_the_bound(gbs, orc):
    """4-bit unsigned x 4-bit unsigned at depth 74000 (K x 15 x 15 just
    below 2^24), all values at their maximum and random: every K piece's
    fp32 sums stay exact integers and the pieces add up exactly."""
    rng = np.random.default_rng(43)
    K = 74000
    probs = []
    for M, N, full in ((130, 40, True), (64, 9, False)):
        lhs = np.full((M, K), 15, np.int8) if full else W.values(rng, (M, K), 4, False)
        rhs = np.full((N, K), 15, np.int8) if full else W.values(rng, (N, K), 4, False)
        probs.append(dict(M=M, K=K, N=N, lb=4, rb=4, ls=False, rs=False, lhs=lhs, rhs=rhs))
    items = build(gbs, probs)
    gbs.batch_set_planes(gbs.PLANES_ALL)
    try:
        b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    finally:
        gbs.batch_set_planes(gbs.PLANES_NONE)
    assert all(b.plan_of(k)[0] == gbs.BATCH_PLANES for k in range(len(probs)))
    zero(items)
    b.run()
    gbs.sync()
    for g, p in zip(results(items, probs), probs):
        assert np.array_equal(g, exact(p))
    assert results(items, probs)[0][0, 0] == K * 225
    b.close()


def test_batch_planes_skinny_mode(gbs, orc):
    """By default every product runs from copies; in the skinny mode
    products of at most 16 rhs rows whose operands do not both fit FP4
    codes read their planes, wider ones and FP4-format ones run from
    copies."""
    assert gbs.batch_get_planes() == gbs.PLANES_NONE
    rng = np.random.default_rng(47)
    shapes = [(200, 3000, 16, 4, 2, False, True),   # planes: 4u lhs, N = 16
              (200, 3000, 17, 4, 2, False, True),   # FP6: N > 16
              (200, 3000, 8, 2, 2, False, False),   # FP4: both fit FP4
              (150, 2000, 3, 1, 4, True, True)]     # planes: bipolar x 4s
    probs = [dict(M=M, K=K, N=N, lb=lb, rb=rb, ls=ls, rs=rs,
                  lhs=W.values(rng, (M, K), lb, ls), rhs=W.values(rng, (N, K), rb, rs))
             for (M, K, N, lb, rb, ls, rs) in shapes]
    items = build(gbs, probs)
    b0 = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    assert [b0.plan_of(k)[0] for k in range(len(probs))] == [2, 2, 1, 2]
    b0.close()
    gbs.batch_set_planes(gbs.PLANES_SKINNY)
    try:
        b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    finally:
        gbs.batch_set_planes(gbs.PLANES_NONE)
    assert [b.plan_of(k)[0] for k in range(len(probs))] == [gbs.BATCH_PLANES, 2, 1,
                                                             gbs.BATCH_PLANES]
    zero(items)
    b.run()
    gbs.sync()
    for g, p in zip(results(items, probs), probs):
        assert np.array_equal(g, exact(p))
    b.close()


def test_batch_refreshes_many_copies_in_one_launch(gbs, orc):
    """Every rhs re-imported before each run (new activations), among them
    one operand shared by an FP4, an FP6 and an int8 product: the run
    rebuilds every stale copy (each (matrix, format) once, in one launch)
    and every result equals the exact product of the new values; a run
    with only two stale copies (rebuilt one launch each) too."""
    gbs.batch_set_planes(gbs.PLANES_NONE)
    try:
        rng = np.random.default_rng(53)
        probs = W.batched_uneven(count=40, seed=17, scale=0.2)
        shared = W.values(rng, (30, 1280), 2, False)
        for lb, ls in ((1, False), (4, False), (7, True)):  # FP4, FP6, int8 with a 2u rhs
            probs.append(dict(M=70, K=1280, N=30, lb=lb, rb=2, ls=ls, rs=False,
                              lhs=W.values(rng, (70, 1280), lb, ls), rhs=shared))
        items = build(gbs, probs)
        # the three last products share one rhs matrix
        for k in (-2, -1):
            items[k][1].close()
            items[k] = (items[k][0], items[-3][1], items[k][2])
        b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
        assert [b.plan_of(len(probs) + k)[0] for k in (-3, -2, -1)] == [1, 2, 0]
        for run in (b.run, b.run_graph, b.run):
            for k, ((l, r, _), p) in enumerate(zip(items, probs)):
                if k >= len(probs) - 2:
                    continue  # shares the rhs imported for product -3
                p["rhs"] = W.values(rng, (p["N"], p["K"]), p["rb"], p["rs"])
                r.import_regular(p["rhs"])
            for p in probs[-2:]:
                p["rhs"] = probs[-3]["rhs"]
            zero(items)
            gbs.sync()
            run()
            gbs.sync()
            for g, p in zip(results(items, probs), probs):
                assert np.array_equal(g, exact(p)), (run.__name__, p["lb"], p["rb"], p["N"])
        # two stale copies (the lhs of the FP4 product and of the FP6 one)
        for k in (-3, -2):
            l, r, _ = items[k]
            p = probs[k]
            p["lhs"] = W.values(rng, (p["M"], p["K"]), p["lb"], p["ls"])
            l.import_regular(p["lhs"])
        zero(items)
        gbs.sync()
        b.run()
        gbs.sync()
        for g, p in zip(results(items, probs), probs):
            assert np.array_equal(g, exact(p))
        b.close()
    finally:
        gbs.batch_set_planes(gbs.PLANES_NONE)


def test_batch_edge_shapes(gbs, orc, planes):
    """One-row operands, depth 1, depths one short of and one past a
    super-step, and last tiles that hold a single row (129 lhs rows, 65 and
    33 rhs rows), under every planes-group mode: each result equals the
    exact product."""
    rng = np.random.default_rng(59)
    shapes = [(1, 1, 1, 2, False, 3, False), (1, 1023, 5, 4, True, 2, True),
              (129, 1025, 65, 3, False, 4, False), (129, 64, 33, 1, True, 3, True),
              (17, 4096, 1, 2, True, 1, True), (300, 2048, 2, 4, False, 1, False)]
    probs = [dict(M=M, K=K, N=N, lb=lb, rb=rb, ls=ls, rs=rs,
                  lhs=W.values(rng, (M, K), lb, ls), rhs=W.values(rng, (N, K), rb, rs))
             for (M, K, N, lb, ls, rb, rs) in shapes]
    items = build(gbs, probs)
    b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    if planes == gbs.PLANES_ALL:
        assert all(b.plan_of(k)[0] == gbs.BATCH_PLANES for k in range(len(probs)))
    for run in (b.run, b.run_graph):
        zero(items)
        gbs.sync()
        run()
        gbs.sync()
        for g, p in zip(results(items, probs), probs):
            assert np.array_equal(g, exact(p)), (run.__name__, p["M"], p["K"], p["N"])
    b.close()


@pytest.mark.parametrize("seed", [61, 62, 63])
def test_batch_planes_group_random_sweep(gbs, orc, seed):
    """Seeded random batches of 1..4-bit products (random signedness, rows
    1..400, depths 1..6000, rhs rows 1..120) with every product the planes
    group holds in it: plain and graph-replayed runs equal the exact
    products."""
    rng = np.random.default_rng(seed)
    probs = []
    for _ in range(30):
        M, N = int(rng.integers(1, 400)), int(rng.integers(1, 120))
        K = int(rng.choice([int(rng.integers(1, 1100)), int(rng.integers(1000, 6000))]))
        lb, rb = int(rng.integers(1, 5)), int(rng.integers(1, 5))
        ls, rs = bool(rng.integers(0, 2)), bool(rng.integers(0, 2))
        probs.append(dict(M=M, K=K, N=N, lb=lb, rb=rb, ls=ls, rs=rs,
                          lhs=W.values(rng, (M, K), lb, ls), rhs=W.values(rng, (N, K), rb, rs)))
    items = build(gbs, probs)
    gbs.batch_set_planes(gbs.PLANES_ALL)
    try:
        b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    finally:
        gbs.batch_set_planes(gbs.PLANES_NONE)
    want = [exact(p) for p in probs]
    for run in (b.run, b.run_graph, b.run_graph):
        zero(items)
        gbs.sync()
        run()
        gbs.sync()
        for g, w, p in zip(results(items, probs), want, probs):
            assert np.array_equal(g, w), (seed, p["M"], p["K"], p["N"], p["lb"], p["ls"],
                                          p["rb"], p["rs"])
    b.close()
