"""bench.py's reporting contract, on the CPU: the one-GPU refusal, the
roofline objects (keys, fraction = achieved / peak, the bound the work
model names), the algorithmic work and byte models, and the committed
traffic summary the headline line reports."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("algo", ["mfma", "mfma_fp4", "mfma_fp6", "mfma_bits", "popcount"])
def test_roofline_objects(algo):
    M, K, N = 4096, 8192, 512
    r = bench.roofline_of(algo, M, K, N, 1, 1, 20e-6)
    assert set(r) == {"bound", "achieved", "peak", "unit", "frac"}
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"], rel=2e-3)
    if algo.startswith("mfma"):
        assert r["bound"] == "mfma"
        assert r["achieved"] == pytest.approx(2 * M * K * N / 20e-6 / 1e12, rel=1e-3)
        assert r["peak"] == (5000.0 if algo == "mfma" else 10100.0)
    else:
        assert r["bound"] in ("valu", "hbm")


def test_popcount_roofline_picks_the_binding_resource():
    # a GEMV over a huge lhs is bound by bytes, a fat product by lane-ops
    skinny = bench.roofline_of("popcount", 32768, 32768, 1, 2, 2, 44e-6)
    fat = bench.roofline_of("popcount", 256, 2304, 50176, 2, 3, 370e-6)
    assert skinny["bound"] == "hbm" and skinny["unit"] == "GB/s"
    assert skinny["achieved"] == pytest.approx(
        bench.plane_bytes(32768, 32768, 1, 2, 2) / 44e-6 / 1e9, rel=1e-3)
    assert fat["bound"] == "valu"
    assert fat["achieved"] == pytest.approx(256 * 50176 * 72 * 2 * 6 / 370e-6 / 1e12, rel=1e-3)


def test_work_and_byte_models():
    # tiled kernel: AND + v_bcnt per 32-bit word and plane pair
    assert bench.valu_lane_ops(256, 2304, 50176, 2, 3) == 256 * 50176 * 72 * 2 * 6
    # skinny kernel (N <= 8, <= 4 planes): 4 AND + 4 v_bcnt + 1 mad per 128-bit slice
    assert bench.valu_lane_ops(1000, 1000, 8, 2, 2) == 1000 * 8 * 8 * 9 * 4
    assert bench.plane_bytes(10, 128, 2, 3, 1) == (10 * 128 * 3 + 2 * 128) // 8 + 4 * 10 * 2
    assert bench.operand_bytes("mfma_fp4", 10, 128, 2, 1, 1) == 12 * 128 * 0.5 + 80
    assert bench.operand_bytes("mfma_fp6", 10, 128, 2, 1, 1) == 12 * 128 * 0.75 + 80
    assert bench.operand_bytes("mfma", 10, 128, 2, 1, 1) == 12 * 128 + 80
    assert bench.operand_bytes("popcount", 10, 128, 2, 3, 1) == bench.plane_bytes(10, 128, 2, 3, 1)
    # the 1-bit matrix-core engine reads the planes, not copies
    assert bench.operand_bytes("mfma_bits", 10, 128, 2, 1, 1) == bench.plane_bytes(10, 128, 2, 1, 1)
    assert bench.ENGINE_DTYPE["mfma_bits"] == "fp4"


@pytest.mark.parametrize("engine", ["mfma_fp4", "mfma_bits"])
def test_headline_traffic_comes_from_the_committed_counter_summary(engine):
    t = bench.measured_traffic(engine)
    assert isinstance(t, int) and t > 0
    kv = {k.strip(): v.strip() for k, v in (ln.split("=", 1) for ln in open(bench.TRAFFIC_FILES[engine])
                                            if "=" in ln and not ln.lstrip().startswith("#"))}
    fetch = float(kv["fetch_kib_per_dispatch"]) * float(kv["fetch_scale"]) * 1024
    write = float(kv["write_kib_per_dispatch"]) * float(kv["write_scale"]) * 1024
    # the summary lists the copy scales rounded to 3 decimals
    assert t == pytest.approx(fetch + write, rel=1e-3)


def test_traffic_is_null_for_engines_not_measured_on_the_headline():
    assert bench.measured_traffic("mfma") is None
    assert bench.measured_traffic("popcount") is None
