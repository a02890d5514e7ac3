"""Benchmark of the MI355X bit-serial GEMM (libgbs_hip.so).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--configs c0,c1,c2]
                    [--engines auto|all] [--no-cpu]

Headline workload (value): the binary layers -- lhs 4096 x 8192 1-bit, rhs
512 x 8192 1-bit, int32 result -- run once bipolar {-1,+1} (Bernoulli 0.5)
and once unsigned {0,1} (30 % ones) per step.  Inputs are packed into
bit-planes in HBM before timing; a step is one gemmBitSerial per encoding
with the result left in HBM; value = effective TOPS = sum of 2*M*K*N over
the step / device time per step (HIP events on the stream the kernels run
on).  The other configurations are reported beside it under "configs":
  c0  plain GEMM 256 x 2048 x 256, lhs 2-bit unsigned, rhs 2-bit signed
  c2  conv-as-GEMM: weights 256 x 2304 2-bit signed, activations
      50176 x 2304 3-bit unsigned (50 % zeros), popcount vs MFMA engine,
      repeated at 4 x 4 bits for the crossover
  c3  512 uneven products (M 64..4096, K 256..8192, N 1..256 log-uniform,
      1..4-bit operands, random signedness) in one persistent launch,
      replayed from a hipGraph, against one launch per product
  c4  serving: 8-layer MLP, 4-bit signed weights resident in HBM, Poisson
      arrivals of 1..64-row 4-bit requests, continuous batching; p50/p99
      latency and TOPS
  c5  the CaffeNet GEMM list of the reference benchmark (8 products, 2-bit
      unsigned x 2-bit unsigned) in one batched launch, against one launch
      per product and the reference CPU loop
  c6  skinny products: lhs 32768 x 32768 2-bit signed (268 MB of planes,
      more than the 256 MB last-level cache, so every pass reads HBM) times
      1 and 8 rhs rows of 2-bit unsigned values (not in the default list)
Every GPU result is checked bit-exact against an exact product before it is
reported.  The reference CPU path (gemmBitSerial compiled from the
reference sources, oracle/_ref, single thread) is timed on a bounded sample
of each configuration.  Prints one JSON line on stdout; progress on stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("bit-exact effective TOPS (2·M·K·N/s) per shape; "
          "batched vs per-call; p99 ms")
SEED0 = 0x67627300
VALU_PEAK = 256 * 64 * 2.4e9      # int32 lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz
HBM_PEAK = 8.0e12                 # B/s
MFMA_I8_PEAK = 5.0e15             # dense int8 ops/s
MFMA_F4_PEAK = 10.1e15            # dense FP4 / FP6 ops/s
# HBM bytes per launch of the headline workload's kernel, per engine that
# runs it (the 1-bit planes kernel AUTO takes, the FP4-copy kernel), measured
# with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE, scaled by a copy of known size)
TRAFFIC_FILES = {"mfma_fp4": os.path.join(ROOT, "profiles", "mfma_fp4_binary_layers_traffic.txt"),
                 "mfma_bits": os.path.join(ROOT, "profiles", "mfma_bits_binary_layers_traffic.txt"),
                 "c3": os.path.join(ROOT, "profiles", "batched_uneven_512_traffic.txt")}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and logical CPU count."""
    name = "unknown CPU"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                name = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return f"{name}, {os.cpu_count()} logical CPUs on the host"


# ------------------------------------------------------------- workloads
from gemmbitserial_amd import workloads as WL  # noqa: E402

CONFIGS = {"c0": ("plain GEMM 256x2048x256, lhs 2-bit unsigned, rhs 2-bit signed",
                  WL.plain_gemm),
           "c1": ("binary layers: lhs 4096x8192 1-bit, rhs 512x8192 1-bit, bipolar "
                  "(Bernoulli 0.5) + unsigned {0,1} (30% ones)", WL.binary_layers),
           "c2": ("conv-as-GEMM: weights 256x2304, activations 50176x2304 (16x56x56 im2col, "
                  "50% zeros), 2s x 3u and 4s x 4u", WL.conv_gemm),
           "c6": ("skinny: lhs 32768x32768 2-bit signed (268 MB of planes), rhs 1 and 8 rows "
                  "2-bit unsigned", WL.skinny)}


def valu_lane_ops(M, K, N, lb, rb):
    """AND/XOR + v_bcnt per 32-bit word per plane pair (tiled kernel); for
    products of at most 8 rhs rows and 4 planes a side (the skinny kernel)
    4 AND + 4 v_bcnt + 1 multiply-add per 128-bit slice per plane pair."""
    if N <= 8 and lb <= 4 and rb <= 4:
        return M * N * ((K + 127) // 128) * 9 * lb * rb
    return M * N * ((K + 31) // 32) * 2 * lb * rb


def plane_bytes(M, K, N, lb, rb):
    return (M * K * lb + N * K * rb) // 8 + 4 * M * N


def int8_bytes(M, K, N):
    return (M + N) * K + 4 * M * N


def operand_bytes(algo, M, K, N, lb, rb):
    """HBM bytes one call reads and writes at least: the bit-planes for the
    popcount engine, the value copies for the matrix cores, plus the result."""
    if algo in ("popcount", "mfma_bits", "mfma_planes"):
        return plane_bytes(M, K, N, lb, rb)
    return (M + N) * K * {"mfma": 1.0, "mfma_fp4": 0.5, "mfma_fp6": 0.75}[algo] + 4 * M * N


def exact_product(O, a, b, lb, rb, ls, rs):
    """The exact int32 result, 2048 lhs rows at a time (float64 BLAS on
    small integers is exact)."""
    bv = O.values_of(b, rb, rs)
    parts = [O.blas_gemm_exact(O.values_of(a[i:i + 2048], lb, ls), bv)
             for i in range(0, a.shape[0], 2048)]
    return O.wrap_i32(np.concatenate(parts, axis=1))


# arithmetic each engine computes in: popcount on 32-bit plane words, int8
# products, or FP4 / FP6 codes of the integer values with fp32 accumulation
ENGINE_DTYPE = {"popcount": "u32", "mfma": "i8", "mfma_fp4": "fp4", "mfma_fp6": "fp6",
                "mfma_bits": "fp4", "mfma_planes": "fp4"}


def measured_traffic(engine):
    """Bytes per launch from the committed counter summary of the headline
    workload's kernel for this engine (the FP4-copy kernel or the 1-bit
    planes kernel), else None."""
    path = TRAFFIC_FILES.get(engine)
    if path is None or not os.path.exists(path):
        return None
    kv = {}
    for ln in open(path):
        if "=" in ln and not ln.lstrip().startswith("#"):
            k, v = ln.split("=", 1)
            kv[k.strip()] = v.strip()
    return int(kv["traffic_bytes_per_launch"]) if "traffic_bytes_per_launch" in kv else None


def roofline_of(algo, M, K, N, lb, rb, sec):
    if algo in ("mfma", "mfma_fp4", "mfma_fp6", "mfma_bits", "mfma_planes"):
        ach = 2.0 * M * K * N / sec
        peak, unit = ((MFMA_I8_PEAK, "TOP/s (int8)") if algo == "mfma"
                      else (MFMA_F4_PEAK, "TOP/s (fp4 operands)")
                      if algo in ("mfma_bits", "mfma_planes")
                      else (MFMA_F4_PEAK, f"TOP/s ({algo[-3:]} operands)"))
        return {"bound": "mfma", "achieved": round(ach / 1e12, 2),
                "peak": peak / 1e12, "unit": unit, "frac": round(ach / peak, 4)}
    ach = valu_lane_ops(M, K, N, lb, rb) / sec
    hbm = plane_bytes(M, K, N, lb, rb) / sec
    if hbm / HBM_PEAK > ach / VALU_PEAK:
        return {"bound": "hbm", "achieved": round(hbm / 1e9, 1), "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": round(hbm / HBM_PEAK, 4)}
    return {"bound": "valu", "achieved": round(ach / 1e12, 3), "peak": round(VALU_PEAK / 1e12, 3),
            "unit": "Tlane-op/s (int32 VALU)", "frac": round(ach / VALU_PEAK, 4)}


# ---------------------------------------------------------------- timing
def time_gemm(gbs, ws, l, r, d_res, algo, steps, warmup, stream):
    for _ in range(warmup):
        ws.gemm_async(l, r, d_res.ptr, algo, stream)
    gbs.sync(stream)
    e0, e1 = gbs.Event(), gbs.Event()
    e0.record(stream)
    for _ in range(steps):
        ws.gemm_async(l, r, d_res.ptr, algo, stream)
    e1.record(stream)
    gbs.sync(stream)
    return e0.elapsed_ms(e1) / steps


def time_dropin(gbs, l, r, M, N, reps):
    """The drop-in call end to end, as gemmBitSerial makes it: one
    synchronous gbs_gemm (AUTO engine) into pinned host memory (the drop-in
    header allocates res with gbs_host_alloc) -- launch, kernels, result
    copy over PCIe and the wait; host clock.  Returns (seconds, result)."""
    import ctypes as C
    p = C.c_void_p()
    gbs.check(gbs.lib().gbs_host_alloc(4 * M * N, C.byref(p)))
    try:
        for _ in range(2):
            gbs.check(gbs.lib().gbs_gemm(None, l.h, r.h, p, 0, gbs.AUTO))
        t0 = time.perf_counter()
        for _ in range(reps):
            gbs.check(gbs.lib().gbs_gemm(None, l.h, r.h, p, 0, gbs.AUTO))
        dt = (time.perf_counter() - t0) / reps
        got = np.ctypeslib.as_array((C.c_int32 * (M * N)).from_address(p.value))
        got = got.reshape(N, M).copy()
    finally:
        gbs.lib().gbs_host_free(p)
    return dt, got


def run_config(gbs, O, name, runs, engines, steps, warmup, stream, ws, cpu_secs):
    out = []
    for enc, a, b, lb, rb, ls, rs in runs:
        M, K = a.shape
        N = b.shape[0]
        log(f"[bench] {name} {enc}: pack lhs {M}x{K} ({lb}b), rhs {N}x{K} ({rb}b)")
        l = gbs.from_values(a, lb, ls)
        r = gbs.from_values(b, rb, rs)
        ws.reserve(l, r)
        d_res = gbs.DeviceBuffer(4 * M * N)
        auto = gbs.select(l, r)
        want = None
        row = {"encoding": enc, "M": M, "K": K, "N": N, "bits": [lb, rb],
               "signed": [ls, rs], "auto": auto, "engines": {}}
        for eng in (["popcount", "mfma", "mfma_fp4", "mfma_fp6", "mfma_bits", "mfma_planes"]
                    if engines == "all"
                    else [auto]):
            log(f"[bench] {name} {enc}: {eng} x {steps}")
            try:
                ms = time_gemm(gbs, ws, l, r, d_res, eng, steps, warmup, stream)
            except gbs.GbsError as e:  # an engine that cannot hold this product exactly
                row["engines"][eng] = {"not_exact": str(e).split(": ", 1)[-1]}
                continue
            got = d_res.to_host(np.int32, (N, M))
            if want is None:
                want = exact_product(O, a, b, lb, rb, ls, rs)
            if not np.array_equal(got, want):
                raise SystemExit(f"{name} {enc} {eng}: result differs from the exact product")
            sec = ms * 1e-3
            row["engines"][eng] = {"us": round(ms * 1e3, 2),
                                   "tops": round(2.0 * M * K * N / sec / 1e12, 2),
                                   "roofline": roofline_of(eng, M, K, N, lb, rb, sec),
                                   "operand_gbs": round(operand_bytes(eng, M, K, N, lb, rb)
                                                        / sec / 1e9, 1),
                                   "bit_exact": True}
            log(f"[bench] {name} {enc}: {eng} {ms * 1e3:.1f} us, "
                f"{2.0 * M * K * N / sec / 1e12:.1f} TOPS, bit-exact")
        # fresh rhs planes before every call (a conv layer's new activations):
        # AUTO, and the int8 copy engine that then rebuilds its copy each time
        fresh = {}
        for eng in ("auto", "mfma"):
            def once(e=eng):
                gbs.check(gbs.lib().gbs_matrix_mark_modified(r.h))
                ws.gemm_async(l, r, d_res.ptr, e, stream)
            try:
                once()
                gbs.sync(stream)
                picked = gbs.select(l, r) if eng == "auto" else eng
                e0, e1 = gbs.Event(), gbs.Event()
                e0.record(stream)
                for _ in range(steps):
                    once()
                e1.record(stream)
                gbs.sync(stream)
            except gbs.GbsError:
                continue
            if not np.array_equal(d_res.to_host(np.int32, (N, M)), want):
                raise SystemExit(f"{name} {enc}: fresh-rhs {eng} differs from the exact product")
            fresh[eng] = {"us": round(e0.elapsed_ms(e1) / steps * 1e3, 2), "engine": picked}
        row["fresh_rhs"] = fresh
        log(f"[bench] {name} {enc}: fresh rhs planes every call: " +
            ", ".join(f"{k} ({v['engine']}) {v['us']} us" for k, v in fresh.items()))
        e2e, got = time_dropin(gbs, l, r, M, N, max(3, min(20, steps)))
        if want is not None and not np.array_equal(got, want):
            raise SystemExit(f"{name} {enc}: drop-in call result differs from the exact product")
        row["dropin_call"] = {"us": round(e2e * 1e6, 1), "result_bytes": 4 * M * N,
                              "note": "synchronous gbs_gemm, AUTO, into pinned host memory: "
                                      "launch + kernels + result copy + wait, host clock"}
        log(f"[bench] {name} {enc}: drop-in call end to end {e2e * 1e6:.1f} us "
            f"({4 * M * N / 1e6:.1f} MB result to the host)")
        if cpu_secs > 0 and O.ref_available():
            # very large lhs: the reference runs on its first rows only
            ms_ = max(1, min(M, (1 << 26) // max(K, 1)))
            ns, reps, res = O.ref_bench_gemm(a[:ms_], b, lb, rb, ls, rs, min_secs=cpu_secs,
                                             max_reps=200)
            if not np.array_equal(res, want[:, :ms_]):
                raise SystemExit(f"{name} {enc}: reference CPU result differs")
            row["cpu_reference"] = {"tops": round(2.0 * ms_ * K * N / ns / 1e3, 4),
                                    "ms_per_call": round(ns / 1e6, 3), "timed_calls": reps,
                                    "cores": 1}
            if ms_ < M:
                row["cpu_reference"]["sample"] = f"first {ms_} of {M} lhs rows"
            log(f"[bench] {name} {enc}: reference CPU {ns / 1e6:.1f} ms/call ({reps} calls)")
        for m in (l, r):
            m.close()
        d_res.close()
        out.append(row)
    return out


def bench_host_import(gbs, O, run, log, cpu_secs):
    """importRegular of the conv activations as the drop-in makes it: an
    int8 matrix in host memory to device planes (copy over PCIe + pack),
    host clock around the synchronous call; the reference's importRegular
    (oracle/_ref, one core) on its first rows beside it."""
    enc, a, b, lb, rb, ls, rs = run
    rows, cols = b.shape
    m = gbs.DeviceMatrix(rb, rows, cols, rs)
    m.import_regular(b)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        m.import_regular(b)
    dt = (time.perf_counter() - t0) / reps
    if not np.array_equal(m.planes(), O.import_regular(b, rb, rs).planes):
        raise SystemExit("c2 host import: planes differ from the oracle's")
    m.close()
    out = {"workload": f"{rows} x {cols} int8 activations ({rb}-bit {'signed' if rs else 'unsigned'}) "
                       "from pageable host memory", "ms": round(dt * 1e3, 3),
           "host_bytes": int(b.nbytes), "gb_per_s": round(b.nbytes / dt / 1e9, 1)}
    if cpu_secs > 0 and O.ref_available():
        k = min(rows, 8192)
        t0, n = time.perf_counter(), 0
        while True:
            O.ref_import_regular(b[:k], rb, rs)
            n += 1
            if time.perf_counter() - t0 > min(cpu_secs, 2.0):
                break
        per = (time.perf_counter() - t0) / n
        out["cpu_reference"] = {"ms": round(per * rows / k * 1e3, 3), "cores": 1,
                                "sample": f"first {k} of {rows} rows, time scaled by rows",
                                "cpu": cpu_model()}
    log(f"[bench] c2 import from host: {out['ms']} ms ({out['gb_per_s']} GB/s of int8)" +
        (f", reference CPU {out['cpu_reference']['ms']} ms" if "cpu_reference" in out else ""))
    return out


def bench_im2col_pack(gbs, steps, warmup, log):
    """Conv activations of the c2 shape (16 x 256 x 56 x 56, 3 x 3, pad 1,
    3-bit unsigned) packed into bit-planes two ways, sources resident in HBM:
    straight from the NCHW tensor (gbs_import_im2col) and from the int8
    im2col matrix (gbs_import).  Both must give the same words."""
    import ctypes as C
    rng = np.random.default_rng(SEED0 + 2)
    x = rng.integers(1, 8, (16, 256, 56, 56)).astype(np.int8)
    x[rng.random(x.shape) < 0.5] = 0
    g = gbs.conv2d(16, 256, 56, 56, 3, 3, 1, 1)
    rows, cols, oh, ow = gbs.conv2d_shape(g)
    xp = np.zeros((16, 256, 58, 58), np.int8)
    xp[:, :, 1:57, 1:57] = x
    host = np.empty((16, oh, ow, 256, 3, 3), np.int8)
    for ky in range(3):
        for kx in range(3):
            host[:, :, :, :, ky, kx] = xp[:, :, ky:ky + oh, kx:kx + ow].transpose(0, 2, 3, 1)
    host = host.reshape(rows, cols)
    d_x, d_m = gbs.DeviceBuffer(x.nbytes), gbs.DeviceBuffer(host.nbytes)
    d_x.copy_from_host(x)
    d_m.copy_from_host(host)
    a, b = gbs.DeviceMatrix(3, rows, cols, False), gbs.DeviceMatrix(3, rows, cols, False)
    st = gbs.stream()

    def timed(fn):
        for _ in range(warmup):
            fn()
        e0, e1 = gbs.Event(), gbs.Event()
        e0.record(st)
        for _ in range(steps):
            fn()
        e1.record(st)
        gbs.sync(st)
        return e0.elapsed_ms(e1) / steps * 1e3
    t_nchw = timed(lambda: gbs.check(gbs.lib().gbs_import_im2col(
        a.h, C.c_void_p(d_x.ptr), gbs.I8, C.byref(g), 1, None)))
    t_mat = timed(lambda: gbs.check(gbs.lib().gbs_import(
        b.h, C.c_void_p(d_m.ptr), gbs.I8, 0, 1, None)))
    if not np.array_equal(a.planes(), b.planes()):
        raise SystemExit("c2 im2col pack: planes differ from the im2col-matrix import")
    out = {"workload": "16x256x56x56 NCHW, 3x3 pad 1, 3-bit unsigned -> 50176 x 2304 planes",
           "from_nchw_us": round(t_nchw, 1), "from_im2col_matrix_us": round(t_mat, 1),
           "nchw_bytes": x.nbytes, "im2col_matrix_bytes": host.nbytes,
           "plane_bytes": 3 * rows * cols // 8, "bit_exact": True}
    log(f"[bench] c2 pack: from NCHW {t_nchw:.1f} us, from the int8 im2col matrix "
        f"{t_mat:.1f} us (identical planes)")
    for m in (a, b, d_x, d_m):
        m.close()
    return out


def bench_batched(gbs, O, steps, warmup, stream, log, name="c3", cpu_secs=0.0):
    from gemmbitserial_amd import workloads as W
    probs = W.batched_uneven() if name == "c3" else W.caffenet()
    log(f"[bench] {name}: generate and pack {len(probs)} products")
    items = []
    for p in probs:
        l = gbs.from_values(p["lhs"], p["lb"], p["ls"])
        r = gbs.from_values(p["rhs"], p["rb"], p["rs"])
        items.append((l, r, gbs.DeviceBuffer(4 * p["M"] * p["N"])))
    b = gbs.Batch([(l, r, d.ptr) for l, r, d in items])
    info = b.info()
    ops = sum(2.0 * p["M"] * p["K"] * p["N"] for p in probs)

    def timed(fn, n):
        for _ in range(warmup):
            fn(stream)
        gbs.sync(stream)
        e0, e1 = gbs.Event(), gbs.Event()
        e0.record(stream)
        for _ in range(n):
            fn(stream)
        e1.record(stream)
        gbs.sync(stream)
        return e0.elapsed_ms(e1) / n

    out = {"workload": ("512 uneven products in one launch" if name == "c3" else
                        "CaffeNet GEMM list, 2u x 2u, in one launch"), "problems": len(probs),
           "tiles": info["tiles"], "persistent_products": info["persistent"],
           "separate_products": info["separate"], "ops": ops}
    cname = name
    for name, fn, n in [("graph", b.run_graph, steps), ("persistent", b.run, steps),
                        ("per_call", b.run_per_call, max(3, steps // 5))]:
        log(f"[bench] {cname}: {name} x {n}")
        ms = timed(fn, n)
        bad = 0
        for (l, r, d), p in zip(items, probs):
            got = d.to_host(np.int32, (p["N"], p["M"]))
            want = O.wrap_i32(O.blas_gemm_exact(O.values_of(p["lhs"], p["lb"], p["ls"]),
                                                O.values_of(p["rhs"], p["rb"], p["rs"])))
            bad += not np.array_equal(got, want)
            if name != "graph":
                break  # full check once; spot-check the other runs
        if bad:
            raise SystemExit(f"{cname} {name}: {bad} products differ from the exact result")
        out[name] = {"ms": round(ms, 4), "tops": round(ops / (ms * 1e-3) / 1e12, 2)}
        log(f"[bench] {cname}: {name} {ms * 1e3:.1f} us, {ops / (ms * 1e-3) / 1e12:.1f} TOPS")
        d0 = items[0][2]
        d0.fill_zero()
    out["speedup_graph_vs_per_call"] = round(out["per_call"]["ms"] / out["graph"]["ms"], 2)
    out["bit_exact"] = True
    # roofline of the persistent launch: it reads every operand's bit-planes
    # and writes every int32 result at least once
    alg = sum(plane_bytes(p["M"], p["K"], p["N"], p["lb"], p["rb"]) for p in probs)
    sec = out["persistent"]["ms"] * 1e-3
    peak_of = {0: MFMA_I8_PEAK, 1: MFMA_F4_PEAK, 2: MFMA_F4_PEAK}
    plans = [b.plan_of(i) for i in range(len(probs))]

    def floor_s(p, grp):
        ops1 = 2.0 * p["M"] * p["K"] * p["N"]
        if grp == gbs.BATCH_PLANES:
            # one FP4 instruction per 2-bit chunk pair: ceil(lb / 2) plane-side
            # chunks x one or two code-side chunks
            cc = 2 if (p["rb"] >= 3 and not p["rs"]) or p["rb"] == 4 else 1
            return ops1 * ((p["lb"] + 1) // 2) * cc / MFMA_F4_PEAK
        return ops1 / peak_of[grp]
    mfma_floor = sum(floor_s(p, pl[0]) for p, pl in zip(probs, plans) if pl[0] >= 0)
    out["roofline"] = {"bound": "hbm", "achieved": round(alg / sec / 1e9, 1),
                       "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                       "frac": round(alg / sec / HBM_PEAK, 4), "algorithmic_bytes": alg,
                       "hbm_floor_us": round(alg / HBM_PEAK * 1e6, 1),
                       "mfma_floor_us": round(mfma_floor * 1e6, 1),
                       "traffic": measured_traffic(cname)}
    copy_plans = [(p, pl) for p, pl in zip(probs, plans) if pl[0] in (0, 1, 2)]
    tabled = sum(1 for p, pl in copy_plans
                 if gbs.batch_tune_get(pl[0], gbs.batch_tune_class_of(p["N"], p["K"]))[0] == pl[1])
    out["tiles_from_table"] = f"{tabled} of {len(copy_plans)} copy-format products"
    out["formats"] = {f: sum(1 for pl in plans if pl[0] == i)
                      for i, f in enumerate(("int8", "fp4", "fp6", "planes"))}
    # the planes group (products read from their bit-planes) against value
    # copies, with static operands and with every rhs re-imported before
    # each run (new activations: copies are rebuilt first, the planes group
    # reads the new planes as they are)
    mode = gbs.batch_get_planes()
    alt = {}
    try:
        for label, m in (("copies", gbs.PLANES_NONE), ("planes_all", gbs.PLANES_ALL)):
            gbs.batch_set_planes(m)
            bm = gbs.Batch([(l, r, d.ptr) for l, r, d in items])

            def alone_timed(n, fresh, bm=bm):
                # each run timed alone, from just before the call (marking
                # 512 matrices through ctypes takes longer than a run)
                tot = 0.0
                for i in range(warmup + n):
                    for _, r, _ in (items if fresh else ()):
                        gbs.check(gbs.lib().gbs_matrix_mark_modified(r.h))
                    e0, e1 = gbs.Event(), gbs.Event()
                    e0.record(stream)
                    bm.run(stream)
                    e1.record(stream)
                    gbs.sync(stream)
                    if i >= warmup:
                        tot += e0.elapsed_ms(e1)
                return tot / n
            row = {}
            for kind in ("static", "static_alone", "fresh_rhs_alone"):
                log(f"[bench] {cname}: {label} {kind} x {steps}")
                ms = (timed(bm.run, steps) if kind == "static" else
                      alone_timed(steps, kind == "fresh_rhs_alone"))
                (l, r, d), p = items[-1], probs[-1]
                if not np.array_equal(d.to_host(np.int32, (p["N"], p["M"])),
                                      O.wrap_i32(O.blas_gemm_exact(
                                          O.values_of(p["lhs"], p["lb"], p["ls"]),
                                          O.values_of(p["rhs"], p["rb"], p["rs"])))):
                    raise SystemExit(f"{cname} {label} {kind}: result differs")
                row[kind + "_us"] = round(ms * 1e3, 1)
                log(f"[bench] {cname}: {label} {kind} {ms * 1e3:.1f} us")
            alt[label] = row
            bm.close()
    finally:
        gbs.batch_set_planes(mode)
    out["planes_group"] = alt
    if cpu_secs > 0 and O.ref_available():
        # the reference loop of benchmark.cpp:213-220 (every product once per
        # pass), timed on every product of the list or, for the 512-product
        # batch, on every 16th product and scaled by the ops
        stride = 1 if len(probs) <= 64 else 16
        sample = probs[::stride]
        ms, sops = 0.0, 0.0
        for p in sample:
            ns, reps, res = O.ref_bench_gemm(p["lhs"], p["rhs"], p["lb"], p["rb"], p["ls"],
                                             p["rs"], min_secs=cpu_secs / len(sample),
                                             max_reps=200)
            want = O.wrap_i32(O.blas_gemm_exact(O.values_of(p["lhs"], p["lb"], p["ls"]),
                                                O.values_of(p["rhs"], p["rb"], p["rs"])))
            if not np.array_equal(res, want):
                raise SystemExit(f"{cname}: reference CPU result differs")
            ms += ns / 1e6
            sops += 2.0 * p["M"] * p["K"] * p["N"]
        tops = sops / (ms * 1e-3) / 1e12
        out["cpu_reference"] = {"ms_per_pass": round(ops / (tops * 1e12) * 1e3, 3),
                                "tops": round(tops, 4), "cores": 1, "cpu": cpu_model(),
                                "sample": (f"all {len(probs)} products" if stride == 1 else
                                           f"every {stride}th product ({len(sample)} of "
                                           f"{len(probs)}, {sops / ops * 100:.1f} % of the ops), "
                                           "pass time scaled by the ops")}
        log(f"[bench] {cname}: reference CPU {out['cpu_reference']['ms_per_pass']:.1f} ms per pass")
    b.close()
    for l, r, d in items:
        l.close(), r.close(), d.close()
    return out


def serve_run(gbs, srv, reqs, speedup, log, keep=()):
    """Feed requests at their Poisson arrival times divided by speedup (None:
    all at once, the server's capacity) and step until every one finished.
    Rows are taken into one reused buffer and kept for the ids in keep.
    Returns (latencies ms, {id: (rows or None, arrival, done)}, wall s, ops,
    hbm bytes, steps)."""
    count = len(reqs)
    keep = set(keep)
    buf = np.empty((64, srv.h[-1]), np.int32)
    st0 = srv.stats()
    t0 = gbs.clock_ns() + 1_000_000
    arrivals = [t0 + (int(r["t"] / speedup * 1e9) if speedup else 0) for r in reqs]
    nxt, lat, res, pending = 0, [], {}, set()
    while len(res) < count:
        now = gbs.clock_ns()
        while nxt < count and arrivals[nxt] <= now:
            srv.submit(reqs[nxt]["id"], reqs[nxt]["x"], arrivals[nxt])
            pending.add(reqs[nxt]["id"])
            nxt += 1
        if not pending:
            continue  # idle until the next arrival
        if srv.step():
            for i in srv.finished():
                y, arr, done = srv.take(i, out=buf)
                res[i] = (y.copy() if i in keep else None, arr, done)
                lat.append((done - arr) / 1e6)
                pending.discard(i)
    t_end = max(t[2] for t in res.values())
    st = srv.stats()
    return (np.array(lat), res, (t_end - arrivals[0]) / 1e9, st["ops"] - st0["ops"],
            st["hbm_bytes"] - st0["hbm_bytes"], st["steps"] - st0["steps"])


def bench_serving(gbs, O, rate, count, log, cpu_secs=0.0):
    from gemmbitserial_amd import workloads as W
    h, Ws, shifts = W.serving_mlp()
    mats = [gbs.from_values(w, 4, True) for w in Ws]
    srv = gbs.Server(mats, shifts, act_bits=4, max_requests=64, max_rows=64)
    # warm-up requests (not measured)
    warm = W.serving_requests(h[0], 8, rate, seed=W.SEED0 + 55)
    for i, r in enumerate(warm):
        srv.submit(10**6 + i, r["x"])
    pend = {10**6 + i for i in range(len(warm))}
    while pend:
        srv.step()
        for i in list(pend):
            if srv.take(i) is not None:
                pend.discard(i)
    reqs = W.serving_requests(h[0], count, rate)
    out = {"workload": "8-layer MLP serving, continuous batching", "hidden": h,
           "shifts": shifts, "requests": count, "rate_per_s": rate,
           "rows_total": int(sum(r["x"].shape[0] for r in reqs))}
    sweep = []
    for r_ in (rate, 4 * rate, 16 * rate, None):
        log(f"[bench] c4: serving {count} requests at "
            f"{'once (capacity)' if r_ is None else f'{r_:.0f}/s'}, hidden sizes {h}")
        sample = list(range(0, count, max(1, count // 16)))
        lat, res, wall, ops, hbm, steps = serve_run(gbs, srv, reqs, r_ and r_ / rate, log,
                                                    keep=sample)
        for i in sample:
            if not np.array_equal(res[i][0], W.mlp_reference(reqs[i]["x"], Ws, shifts)):
                raise SystemExit(f"c4: request {i} differs from the exact forward pass")
        row = {"rate_per_s": r_ if r_ else "all at once",
               "p50_ms": round(float(np.percentile(lat, 50)), 3),
               "p99_ms": round(float(np.percentile(lat, 99)), 3),
               "tops": round(ops / wall / 1e12, 3), "steps": steps, "wall_s": round(wall, 4),
               "hbm_gbs": round(hbm / wall / 1e9, 1), "bit_exact_sample": len(sample)}
        sweep.append(row)
        log(f"[bench] c4: p50 {row['p50_ms']} ms, p99 {row['p99_ms']} ms, {row['tops']} TOPS, "
            f"{steps} steps")
    # admission by memory: a second server whose budget holds 256 rows of
    # every pool and buffer (the first one's bytes per row), all requests
    # at once -- cohorts are cut by the row capacity, the rest wait
    st1 = srv.stats()
    per_row = st1["bytes_reserved"] // st1["row_capacity"]
    tight = gbs.Server(mats, shifts, act_bits=4, max_requests=64, max_rows=64,
                       mem_budget=per_row * 256)
    log(f"[bench] c4: memory budget of {per_row * 256 / 1e6:.1f} MB "
        f"({tight.stats()['row_capacity']} rows), {count} requests at once")
    sample = list(range(0, count, max(1, count // 16)))
    lat, res, wall, ops, hbm, steps = serve_run(gbs, tight, reqs, None, log, keep=sample)
    for i in sample:
        if not np.array_equal(res[i][0], W.mlp_reference(reqs[i]["x"], Ws, shifts)):
            raise SystemExit(f"c4: request {i} differs under the tight memory budget")
    out["admission_by_memory"] = {
        "mem_budget_bytes": int(per_row * 256), "row_capacity": tight.stats()["row_capacity"],
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "tops": round(ops / wall / 1e12, 3), "steps": steps}
    log(f"[bench] c4: tight budget: p50 {out['admission_by_memory']['p50_ms']} ms, "
        f"{steps} steps")
    tight.close()
    base, cap = sweep[0], sweep[-1]
    out.update({k: base[k] for k in ("p50_ms", "p99_ms", "tops", "steps", "wall_s")})
    out["bit_exact_sample"] = base["bit_exact_sample"]
    out["sweep"] = sweep
    out["capacity_tops"] = cap["tops"]
    st = srv.stats()
    out["row_capacity"] = st["row_capacity"]
    # roofline at capacity: the step launches' algorithmic bytes (weight
    # planes, input and output activation planes, int32 rows out of the last
    # layer) over the wall time of the all-at-once run
    out["roofline"] = {"bound": "hbm", "achieved": cap["hbm_gbs"], "peak": HBM_PEAK / 1e9,
                       "unit": "GB/s", "frac": round(cap["hbm_gbs"] * 1e9 / HBM_PEAK, 4),
                       "at": "capacity (all requests at once), wall clock",
                       "mfma_int8_frac": round(cap["tops"] * 1e12 / MFMA_I8_PEAK, 4),
                       "note": "the step launches run int8 matrix-core GEMMs; at capacity "
                               "the loop is paced by host work per step, not by either roof"}
    if cpu_secs > 0 and O.ref_available():
        # the reference forward pass of sample requests: per layer one
        # gemmBitSerial (weights 4-bit signed x activations 4-bit unsigned),
        # requantized between layers; timed per call like benchmark.cpp:93-111
        ns_tot, ops_tot, picks = 0.0, 0.0, reqs[: 4]
        for r in picks:
            a = r["x"].astype(np.int8)
            for l, w in enumerate(Ws):
                ns, reps, acc = O.ref_bench_gemm(w, a, 4, 4, True, False,
                                                 min_secs=cpu_secs / (len(picks) * len(Ws)),
                                                 max_reps=50)
                ns_tot += ns
                ops_tot += 2.0 * w.shape[0] * w.shape[1] * a.shape[0]
                if l + 1 < len(Ws):
                    a = np.clip(acc.astype(np.int64) >> shifts[l], 0, 15).astype(np.int8)
            if not np.array_equal(acc, W.mlp_reference(r["x"], Ws, shifts)):
                raise SystemExit("c4: reference CPU forward pass differs")
        out["cpu_reference"] = {"tops": round(ops_tot / ns_tot / 1e3, 4),
                                "ms_per_request": round(ns_tot / len(picks) / 1e6, 3),
                                "cores": 1, "cpu": cpu_model(),
                                "sample": f"forward pass of the first {len(picks)} requests "
                                          f"({sum(r['x'].shape[0] for r in picks)} rows)"}
        log(f"[bench] c4: reference CPU {out['cpu_reference']['ms_per_request']} ms per request")
    srv.close()
    for m in mats:
        m.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--configs", default="c1,c0,c2,c3,c4,c5")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--engines", default="all", choices=["auto", "all"])
    ap.add_argument("--cpu-secs", type=float, default=4.0,
                    help="reference CPU timing budget per run (0 = skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="BYTES",
                    help="also run device-to-device copies of BYTES (a kernel of known "
                         "HBM traffic, for scaling profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build "
              "has no multi-GPU path: every problem here fits one MI355X.", file=sys.stderr)
        sys.exit(2)
    cpu_secs = 0.0 if args.no_cpu else args.cpu_secs

    import gemmbitserial_amd as gbs
    from oracle import orc as O
    gbs.init(0)
    stream = gbs.stream()
    ws = gbs.Workspace()

    if args.copy_calibrate:
        src, dst = gbs.DeviceBuffer(args.copy_calibrate), gbs.DeviceBuffer(args.copy_calibrate)
        src.fill_zero()
        for _ in range(3):
            dst.copy_from(src)
        src.close(), dst.close()
    names = [c.strip() for c in args.configs.split(",") if c.strip()]
    if "c1" not in names:
        names = ["c1"] + names
    results = {}
    for i, name in enumerate(names):
        if name in ("c3", "c5"):
            results[name] = bench_batched(gbs, O, args.steps, args.warmup, stream, log, name,
                                          cpu_secs)
            continue
        if name == "c4":
            results[name] = bench_serving(gbs, O, args.serve_rate, args.serve_requests, log,
                                          cpu_secs)
            continue
        desc, make = CONFIGS[name]
        runs = make(np.random.default_rng(SEED0 + int(name[1])))
        results[name] = {"workload": desc, "seed": SEED0 + int(name[1]),
                         "runs": run_config(gbs, O, name, runs, args.engines, args.steps,
                                            args.warmup, stream, ws, cpu_secs)}
        if name == "c2":
            results[name]["pack"] = bench_im2col_pack(gbs, min(args.steps, 20), 2, log)
            results[name]["import_from_host"] = bench_host_import(gbs, O, runs[0], log,
                                                                  cpu_secs)
    # headline: the binary layers through gemmBitSerial's own engine choice
    c1 = results["c1"]["runs"]
    step_ms = sum(r["engines"][r["auto"]]["us"] for r in c1) / 1e3
    ops = sum(2.0 * r["M"] * r["K"] * r["N"] for r in c1)
    value = ops / (step_ms * 1e-3) / 1e12
    # dominant kernel of the step: the engine AUTO picked (same for both runs)
    eng = c1[0]["auto"]
    r0 = c1[0]
    roof = roofline_of(eng, r0["M"], r0["K"], r0["N"], 1, 1, step_ms * 1e-3 / len(c1))
    roof["traffic"] = measured_traffic(eng)
    cpu = None
    if cpu_secs > 0 and all("cpu_reference" in r for r in c1):
        cpu_ms = sum(r["cpu_reference"]["ms_per_call"] for r in c1)
        cpu = {"value": round(ops / (cpu_ms * 1e-3) / 1e12, 4), "unit": "TOPS", "cores": 1,
               "kind": "reference",
               "sample": "the two full binary-layer GEMMs, " + ", ".join(
                   f"{r['encoding']}: {r['cpu_reference']['timed_calls']} timed calls after 1 "
                   f"warm-up" for r in c1) + f"; {cpu_model()}"}
    line = {"metric": METRIC, "value": round(value, 2), "unit": "TOPS",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(step_ms, 5), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": ENGINE_DTYPE[eng],
            "data": "synthetic (seeded numpy), packed to bit-planes in HBM before timing",
            "config": {"workload": CONFIGS["c1"][0], "engine": eng,
                       "configs": results},
            "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
