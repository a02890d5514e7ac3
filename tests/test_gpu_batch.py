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


def test_batch_mixed_small(gbs, orc):
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


def test_batch_order_does_not_matter(gbs, orc):
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


def test_batch_full_size_c3(gbs, orc):
    """The 512-product uneven workload at full size: every result exact."""
    probs = W.batched_uneven()
    items = build(gbs, probs)
    b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    b.run_graph()
    gbs.sync()
    for g, p in zip(results(items, probs), probs):
        assert np.array_equal(g, exact(p))


def test_batch_caffenet_list(gbs, orc):
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


def test_batch_sees_reimported_operands(gbs, orc):
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
        for f in range(3):
            for c in range(n):
                s, pc = saved[f][c]
                gbs.batch_tune_set(f, c, s, pc if s >= 0 else 16)
        for x in items:
            for y in x:
                y.close()
