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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--reps", type=int, default=10)
    a = ap.parse_args()
    shapes, splits, detail = tune(a.reps, log=lambda *x: print(*x, file=sys.stderr, flush=True))
    write_table(shapes, splits, detail, a.out)
    print(json.dumps({"shapes": shapes, "splits": splits, "out": a.out}))


if __name__ == "__main__":
    main()
