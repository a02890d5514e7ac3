"""Tile autotuner of the matrix-core engine.

Products fall into shape classes (buckets of M, N and K; gbs_tune_class_of).
For every class and operand format (int8, FP4, FP6 copies) this times
each compiled tile shape with each K split (1, 2, 4, ... while a split
keeps at least 256 bytes of K) on a representative product (2-bit signed x
2-bit unsigned, copies current, result in HBM) and keeps the fastest pair;
the table is written as text, one "class shape splits format microseconds"
line per class and format, and loaded by the library at gbs_init when it
sits next to libgbs_hip.so (gemmbitserial_amd/tune_table.txt) or through
gbs_tune_load.  The times let AUTO pick FP6 over int8 copies only for the
classes where FP6 measured faster.

    python -m gemmbitserial_amd.tune [--out PATH] [--reps 20]

The persistent launch has a table of its own (a shared queue needs no grid
filling, so the standalone table's shapes do not carry over): per batch
class (rhs-row bucket x depth bucket, gbs_batch_tune_class_of) and copy
format, the tile shape and the largest K piece of a work item.  --batch
tunes it by coordinate descent on whole persistent launches of a mixed
workload (512 uneven products, 1..4-bit operands, random signedness, the
same generator as the batched benchmark with another seed): for every
class present, heaviest first, each carried shape x K piece of 2, 4 or
8 KiB (nominal) is planned into the batch and the launch timed (best of
three runs of --reps launches); a candidate replaces the class's entry
when the launch gets at least 0.5 % faster.  The result is checked bit
for bit against one launch per product and written to
batch_tune_table.txt, which gbs_init loads.

    python -m gemmbitserial_amd.tune --batch [--out PATH] [--reps 10] [--passes 2]

The serving loop's step launch plans its int8 products from the same
table (format 0).  --serving tunes those entries on the serving workload
itself (the 8-layer 4-bit MLP, Poisson requests of 1..64 rows, another
seed than the benchmark's): for every int8 class its products fall in,
each carried shape x K piece of 1/2, 1, 2 or 4 KiB is planned and the
median latency at 2000 and 8000 requests/s measured; a candidate replaces
the entry when that mean drops by at least 2 %.  The int8 lines of the
table are rewritten, the other formats' lines kept.

    python -m gemmbitserial_amd.tune --serving [--out PATH] [--passes 1]
"""
import argparse
import itertools
import json
import os
import sys

import numpy as np

import gemmbitserial_amd as gbs
from gemmbitserial_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tune_table.txt")
DEFAULT_BATCH_OUT = os.path.join(HERE, "batch_tune_table.txt")
BATCH_PIECES = (8, 16, 32)  # K units of 256 nominal bytes: 2, 4, 8 KiB
# representative sizes of the class buckets (see shape_class in
# csrc/gbs_mfma_tile.h): M <256 / <1024 / <4096 / <16384 / >=16384,
# N <=16 / <=32 / <=64 / <=256 / <=2048 / >2048, K <1024 / <4096 / >=4096.
# The products of the largest buckets are kept to a few hundred MB.
REP_M = [160, 640, 2048, 4096, 32768]
REP_N = [8, 24, 48, 256, 512, 16384]
REP_K = [512, 2048, 8192]
SHAPES = ["128x128", "64x128", "32x256/k64", "16x256", "64x64", "128x64", "128x128/k64",
          "128x256/k64", "128x128/pp", "128x256/pp", "64x128/pp", "256x256/k64",
          "256x256", "64x64/dl4", "64x64/k256/dl3", "128x128/dl3", "128x64/dl3",
          "64x128/dl3", "64x64/dl3", "64x64/dl2x2g", "64x128/dl2x2g", "64x64/dl3x2g"]


FORMATS = [("int8", gbs.MFMA), ("fp4", gbs.MFMA_FP4), ("fp6", gbs.MFMA_FP6)]


def time_config(l, r, d, algo, shape, splits, reps, stream):
    call = gbs.lib().gbs_gemm_async_config
    for _ in range(3):
        gbs.check(call(None, l.h, r.h, d.ptr, algo, shape, splits, stream))
    gbs.sync(stream)
    e0, e1 = gbs.Event(), gbs.Event()
    e0.record(stream)
    for _ in range(reps):
        gbs.check(call(None, l.h, r.h, d.ptr, algo, shape, splits, stream))
    e1.record(stream)
    gbs.sync(stream)
    return e0.elapsed_ms(e1) / reps


def tune(reps=10, log=print):
    gbs.init(0)
    stream = gbs.stream()
    n_classes, n_shapes = gbs.tune_classes()
    rng = np.random.default_rng(W.SEED0 + 99)
    nf = len(FORMATS)
    shapes, splits = [-1] * (nf * n_classes), [0] * (nf * n_classes)
    detail = {}
    for M, N, K in itertools.product(REP_M, REP_N, REP_K):
        if M * N > 2 ** 26:  # 256 MB of result: left to the built-in heuristic
            continue
        cls = gbs.tune_class_of(M, N, K)
        l = gbs.from_values(W.values(rng, (M, K), 2, True), 2, True)
        r = gbs.from_values(W.values(rng, (N, K), 2, False), 2, False)
        ws = gbs.Workspace()
        ws.reserve(l, r)
        d = gbs.DeviceBuffer(4 * M * N)
        # every winner must agree with the unsplit int8 64 x 64 product
        time_config(l, r, d, gbs.MFMA, 4, 1, 1, stream)
        want = d.to_host(np.int32, (N, M))
        for f, (fname, algo) in enumerate(FORMATS):
            kbytes = K if algo == gbs.MFMA else K // 2  # nominal bytes (64 per k-step)
            times = {}
            for sh in range(n_shapes):
                sp = 1
                while sp == 1 or kbytes // sp >= 256:
                    times[(sh, sp)] = time_config(l, r, d, algo, sh, sp, reps, stream)
                    sp *= 2
            (bsh, bsp), bt = min(times.items(), key=lambda kv: kv[1])
            time_config(l, r, d, algo, bsh, bsp, 1, stream)
            if not np.array_equal(d.to_host(np.int32, (N, M)), want):
                raise SystemExit(f"{fname} config {SHAPES[bsh]}/{bsp} disagrees at {M}x{K}x{N}")
            shapes[f * n_classes + cls], splits[f * n_classes + cls] = bsh, bsp
            unsplit = [round(times[(x, 1)] * 1e3, 2) for x in range(n_shapes)]
            detail[(f, cls)] = {"M": M, "N": N, "K": K, "best": f"{SHAPES[bsh]} split {bsp}",
                                "us_best": round(bt * 1e3, 2), "us_unsplit_by_shape": unsplit}
            log(f"{fname:4s} class {cls:2d} M={M:5d} N={N:5d} K={K:5d}: {SHAPES[bsh]:12s} "
                f"x{bsp:<3d} {bt * 1e3:7.1f} us  (unsplit: " +
                " ".join(f"{t:.1f}" for t in unsplit) + ")")
        for x in (l, r, d):
            x.close()
        ws.close()
    return shapes, splits, detail


def write_table(shapes, splits, detail, path):
    n_classes = len(shapes) // len(FORMATS)
    with open(path, "w") as f:
        f.write("# matrix-core tile shape and K split per shape class and operand format, "
                "measured by gemmbitserial_amd.tune\n")
        f.write("# shapes: " + " ".join(f"{i}={s}" for i, s in enumerate(SHAPES)) + "\n")
        f.write("# formats: " + " ".join(f"{i}={n}" for i, (n, _) in enumerate(FORMATS)) + "\n")
        f.write("# class shape splits format us  # representative M N K, unsplit times\n")
        for i, sh in enumerate(shapes):
            if sh >= 0:
                fmt, cls = divmod(i, n_classes)
                d = detail.get((fmt, cls), {})
                f.write(f"{cls} {sh} {splits[i]} {fmt} {d.get('us_best', 0)}  # M={d.get('M')} "
                        f"N={d.get('N')} K={d.get('K')} "
                        f"unsplit={d.get('us_unsplit_by_shape')}\n")


def _plane_bytes(p):
    return (p["M"] * p["K"] * p["lb"] + p["N"] * p["K"] * p["rb"]) // 8 + 4 * p["M"] * p["N"]


def tune_batch(reps=10, passes=2, seed=W.SEED0 + 33, log=print):
    """Coordinate descent over the batch tile table on one mixed workload;
    returns (entries, detail): entries maps (fmt, cls) -> (shape, piece)."""
    gbs.init(0)
    stream = gbs.stream()
    probs = W.batched_uneven(seed=seed)
    mats = []
    for p in probs:
        mats.append((gbs.from_values(p["lhs"], p["lb"], p["ls"]),
                     gbs.from_values(p["rhs"], p["rb"], p["rs"]),
                     gbs.DeviceBuffer(4 * p["M"] * p["N"])))
    problems = [(l, r, d.ptr) for l, r, d in mats]
    n_cls = gbs.batch_tune_classes()
    for f in range(len(FORMATS)):
        for c in range(n_cls):
            gbs.batch_tune_set(f, c, -1)

    def measure():
        b = gbs.Batch(problems)
        for _ in range(3):
            b.run(stream)
        gbs.sync(stream)
        best = float("inf")
        for _ in range(3):
            e0, e1 = gbs.Event(), gbs.Event()
            e0.record(stream)
            for _ in range(reps):
                b.run(stream)
            e1.record(stream)
            gbs.sync(stream)
            best = min(best, e0.elapsed_ms(e1) / reps)
        plans = [b.plan_of(i) for i in range(len(problems))]
        b.close()
        return best * 1e3, plans

    t_builtin, plans = measure()
    log(f"batch: built-in rule {t_builtin:.1f} us")
    weight = {}
    for p, pl in zip(probs, plans):
        if pl[0] >= 0:
            key = (pl[0], gbs.batch_tune_class_of(p["N"], p["K"]))
            weight[key] = weight.get(key, 0) + _plane_bytes(p)
    order = sorted(weight, key=lambda k: -weight[k])
    # start every present class from the shape and piece the built-in rule gave it
    entries = {}
    for p, pl in zip(probs, plans):
        if pl[0] >= 0:
            key = (pl[0], gbs.batch_tune_class_of(p["N"], p["K"]))
            entries.setdefault(key, (pl[1], 16))
    for (f, c), (sh, pc) in entries.items():
        gbs.batch_tune_set(f, c, sh, pc)
    cur, _ = measure()
    for it in range(passes):
        for key in order:
            f, c = key
            best, best_t = entries[key], cur
            for sh in gbs.batch_tune_shapes(f):
                for pc in BATCH_PIECES:
                    if (sh, pc) == entries[key]:
                        continue
                    gbs.batch_tune_set(f, c, sh, pc)
                    t, _ = measure()
                    if t < best_t * 0.995:
                        best, best_t = (sh, pc), t
            entries[key] = best
            gbs.batch_tune_set(f, c, *best)
            cur = best_t
            log(f"batch pass {it}: {FORMATS[f][0]} class {c:2d} ({weight[key] / 1e6:6.1f} MB) "
                f"-> shape {SHAPES[best[0]]} piece {best[1]}: {cur:.1f} us")
        cur, _ = measure()  # re-measure the table as a whole after each pass
        log(f"batch pass {it}: whole table {cur:.1f} us (built-in {t_builtin:.1f} us)")
    # bit-exact against one launch per product
    b = gbs.Batch(problems)
    b.run(stream)
    gbs.sync(stream)
    got = [d.to_host(np.int32, (p["N"], p["M"])) for (l, r, d), p in zip(mats, probs)]
    for _, _, d in mats:
        d.fill_zero()
    b.run_per_call(stream)
    gbs.sync(stream)
    for g, (l, r, d), p in zip(got, mats, probs):
        if not np.array_equal(g, d.to_host(np.int32, (p["N"], p["M"]))):
            raise SystemExit("tuned batch differs from the per-product launches")
    b.close()
    counts = {}
    for p, pl in zip(probs, plans):
        if pl[0] >= 0:
            key = (pl[0], gbs.batch_tune_class_of(p["N"], p["K"]))
            counts[key] = counts.get(key, 0) + 1
    detail = {"workload": f"batched_uneven(seed={seed:#x}), {len(probs)} products",
              "builtin_us": round(t_builtin, 1), "tuned_us": round(cur, 1),
              "products": counts, "mbytes": {k: round(v / 1e6, 1) for k, v in weight.items()}}
    for x in mats:
        for y in x:
            y.close()
    return entries, detail


def _serve_p50(srv, reqs, speedup, h_last):
    """Median latency (ms) of reqs fed at their arrival times / speedup."""
    buf = np.empty((64, h_last), np.int32)
    t0 = gbs.clock_ns() + 1_000_000
    arr = [t0 + int(r["t"] / speedup * 1e9) for r in reqs]
    nxt, lat, pending = 0, [], set()
    while len(lat) < len(reqs):
        now = gbs.clock_ns()
        while nxt < len(reqs) and arr[nxt] <= now:
            srv.submit(reqs[nxt]["id"], reqs[nxt]["x"], arr[nxt])
            pending.add(reqs[nxt]["id"])
            nxt += 1
        if pending and srv.step():
            for i in srv.finished():
                _, a, d = srv.take(i, out=buf)
                lat.append((d - a) / 1e6)
                pending.discard(i)
    return float(np.median(lat))


def tune_serving(passes=1, seed=W.SEED0 + 66, log=print):
    """Coordinate descent over the int8 batch-table entries the serving
    step launch uses; returns (entries, detail)."""
    gbs.init(0)
    h, Ws, shifts = W.serving_mlp()
    mats = [gbs.from_values(w, 4, True) for w in Ws]
    reqs = W.serving_requests(h[0], 300, 2000.0, seed=seed)
    n_cls = gbs.batch_tune_classes()
    for c in range(n_cls):
        gbs.batch_tune_set(0, c, -1)

    def measure():
        srv = gbs.Server(mats, shifts, act_bits=4, max_requests=64, max_rows=64)
        _serve_p50(srv, reqs[:16], 1.0, h[-1])   # warm-up
        t = [_serve_p50(srv, reqs, sp, h[-1]) for sp in (1.0, 4.0)]
        srv.close()
        return sum(t) / len(t)

    # classes of the serving products: a cohort of 1..64 rows per layer at
    # these loads (rhs rows <= 64), depth h_l
    classes = sorted({gbs.batch_tune_class_of(n, k) for n in (8, 24, 48) for k in h[:-1]})
    base = measure()
    log(f"serving: built-in rule {base * 1e3:.1f} us median latency")
    cur, entries = base, {}
    for it in range(passes):
        for c in classes:
            best, best_t = entries.get(c), cur
            for sh in gbs.batch_tune_shapes(0):
                for pc in (2, 4, 8, 16):
                    gbs.batch_tune_set(0, c, sh, pc)
                    t = measure()
                    if t < best_t * 0.98:
                        best, best_t = (sh, pc), t
            if best is None:
                gbs.batch_tune_set(0, c, -1)
            else:
                gbs.batch_tune_set(0, c, *best)
                entries[c] = best
            cur = best_t
            log(f"serving pass {it}: int8 class {c:2d} -> "
                f"{SHAPES[best[0]] + ' piece ' + str(best[1]) if best else 'built-in'}: "
                f"{cur * 1e3:.1f} us")
    final = measure()
    for m in mats:
        m.close()
    return entries, {"builtin_ms": round(base, 4), "tuned_ms": round(final, 4),
                     "workload": f"serving_mlp(), 300 requests (seed {seed:#x}) at 2000 "
                                 "and 8000 requests/s, mean of the two median latencies"}


def write_serving_entries(entries, detail, path):
    """Rewrite the int8 (format 0) lines of a batch table, keeping the rest."""
    keep = []
    if os.path.exists(path):
        for ln in open(path):
            body = ln.split("#", 1)[0].split()
            if ln.startswith("# serving") or (len(body) >= 2 and body[1] == "0"):
                continue
            keep.append(ln.rstrip("\n"))
    with open(path, "w") as f:
        for ln in keep:
            f.write(ln + "\n")
        f.write(f"# serving (int8) entries, measured by gemmbitserial_amd.tune --serving on "
                f"{detail['workload']}: {detail['tuned_ms']} ms (built-in rule "
                f"{detail['builtin_ms']} ms)\n")
        for c, (sh, pc) in sorted(entries.items()):
            f.write(f"{c} 0 {sh} {pc}  # serving step launch\n")


def write_batch_table(entries, detail, path):
    with open(path, "w") as f:
        f.write("# persistent-launch tile shape and K piece per batch class and copy format, "
                "measured by gemmbitserial_amd.tune --batch\n")
        f.write(f"# workload: {detail['workload']}; whole launch {detail['tuned_us']} us "
                f"(built-in rule {detail['builtin_us']} us)\n")
        f.write("# classes: rhs rows <=16/<=32/<=64/<=128/<=256/more x depth <1024/<4096/more "
                "(class = 3 * rows bucket + depth bucket)\n")
        f.write("# shapes: " + " ".join(f"{i}={s}" for i, s in enumerate(SHAPES)) + "\n")
        f.write("# formats: " + " ".join(f"{i}={n}" for i, (n, _) in enumerate(FORMATS)) + "\n")
        f.write("# class format shape piece_units  # products in the workload, plane MB\n")
        for (fmt, cls), (sh, pc) in sorted(entries.items(), key=lambda kv: (kv[0][1], kv[0][0])):
            f.write(f"{cls} {fmt} {sh} {pc}  # {detail['products'].get((fmt, cls), 0)} products, "
                    f"{detail['mbytes'].get((fmt, cls), 0)} MB\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--batch", action="store_true", help="tune the persistent launch's table")
    ap.add_argument("--serving", action="store_true",
                    help="tune the int8 entries the serving step launch uses")
    ap.add_argument("--passes", type=int, default=2)
    a = ap.parse_args()
    if a.serving:
        entries, detail = tune_serving(a.passes if a.passes else 1,
                                       log=lambda *x: print(*x, file=sys.stderr, flush=True))
        out = a.out or DEFAULT_BATCH_OUT
        write_serving_entries(entries, detail, out)
        print(json.dumps({"builtin_ms": detail["builtin_ms"], "tuned_ms": detail["tuned_ms"],
                          "entries": {str(k): v for k, v in entries.items()}, "out": out}))
        return
    if a.batch:
        entries, detail = tune_batch(a.reps, a.passes,
                                     log=lambda *x: print(*x, file=sys.stderr, flush=True))
        out = a.out or DEFAULT_BATCH_OUT
        write_batch_table(entries, detail, out)
        print(json.dumps({"builtin_us": detail["builtin_us"], "tuned_us": detail["tuned_us"],
                          "entries": len(entries), "out": out}))
        return
    a.out = a.out or DEFAULT_OUT
    shapes, splits, detail = tune(a.reps, log=lambda *x: print(*x, file=sys.stderr, flush=True))
    write_table(shapes, splits, detail, a.out)
    print(json.dumps({"shapes": shapes, "splits": splits, "out": a.out}))


if __name__ == "__main__":
    main()
